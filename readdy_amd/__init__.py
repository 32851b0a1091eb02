"""MI355X-native ReaDDy compute kernel.

The HIP engine (libreaddy_amd.so, C ABI in include/readdy_amd.h) implements ReaDDy's per-step
particle pipeline: cell-linked neighbour list, pair and bonded forces, Euler Brownian dynamics
and the UncontrolledApproximation / Gillespie reaction handlers, for one system or thousands of
replicas stepped together.  `Engine` drives it from Python; the C++ host core and the "MI355X" kernel
plugin (readdy_amd/host) carry the reference's Kernel / StateModel / ActionFactory / ObservableFactory
object contract; `readdy_amd.api` offers ReaDDy's Python calls (ReactionDiffusionSystem, simulation(),
observe.*, run) over the engine.
"""
from .engine import Engine, bin_capacity_for, device_info  # noqa: F401

KERNEL_NAME = "MI355X"
