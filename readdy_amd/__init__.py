"""MI355X-native ReaDDy compute kernel.

The HIP engine (libreaddy_amd.so, C ABI in include/readdy_amd.h) implements ReaDDy's per-step
particle pipeline: cell-linked neighbour list, pair and bonded forces, Euler Brownian dynamics
and the UncontrolledApproximation / Gillespie reaction handlers, for one system or thousands of
replicas stepped together.  `readdy_amd.api` mirrors the `readdy` Python API so that
`ReactionDiffusionSystem(...).simulation(kernel="MI355X")` is a drop-in.
"""
from .engine import Engine, bin_capacity_for, device_info  # noqa: F401

KERNEL_NAME = "MI355X"
