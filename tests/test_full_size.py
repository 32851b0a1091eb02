"""Parity at full size: the 10^6-particle A+B<->C workload of the benchmark (periodic 200^3 box, 40/40/20 %
A/B/C, seed 1; harmonic repulsion on all pairs, A-A weak interaction to 2.0, fusion and fission) against
the FP64 oracle running on the host cores: the neighbour pair set within the 2.0 cutoff and the reaction
candidate set (before the Bernoulli draws) bit-exact, forces and potential energy within 1e-5 relative,
on the initial state and on the state after 10 graph-replayed fused steps (the state the benchmark
measures; particles compared by handle).
"""
import os

import numpy as np
import pytest

from oracle import pyoracle as po

from readdy_amd import workloads as W

pytestmark = pytest.mark.gpu

L = pytest.importorskip("readdy_amd._lib")
from readdy_amd import Engine  # noqa: E402

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
BOX = 200.0
DIFF = [1.0, 1.0, 0.5]


def _sorted_rows(a):
    a = np.asarray(a, dtype=np.uint64)
    return a[np.lexsort(a.T[::-1])]


@pytest.fixture(scope="module")
def c1():
    print("building the 10^6-particle system", flush=True)
    pos, types = W.c1_system()
    pots, rxs = W.c1_chemistry(L)
    e = Engine(3, DIFF, potentials=pots, reactions=rxs)
    r = e.add_replica([BOX] * 3, (1, 1, 1), seed=1, max_particles=2 * len(types))
    e.set_particles(r, pos, types)
    return e, r, pots, rxs


def _oracle(pots, rxs, pos, types, ids):
    o = po.System(3, [BOX] * 3, [1, 1, 1], diffusion=DIFF)
    for kind, a, b, p in pots:
        o.add_pair_potential(kind, a, b, p)
    for x in rxs:
        o.add_reaction(x["kind"], x["educts"], x["products"], x["rate"], x.get("educt_distance", 0.0),
                       x.get("product_distance", 0.0))
    o.add_particles(pos, types.astype(np.int32), ids=ids)
    o.set_threads(THREADS)
    o.nl_setup(o.max_cutoff())
    return o


@pytest.mark.parametrize("steps", [0, 10])
def test_c1_full_size_parity(c1, steps):
    e, r, pots, rxs = c1
    if steps:
        print(f"running {steps} fused steps", flush=True)
        e.run(L.HANDLER_UA, 0.005, steps, graph=True)
    pos, types, hs = e.get_particles(r)
    assert len(types) > 990_000
    print(f"oracle on {THREADS} threads, {len(types)} particles", flush=True)
    o = _oracle(pots, rxs, pos, types, hs.astype(np.int64))
    # neighbour pair set within the largest cutoff: bit-exact
    gpu_pairs = _sorted_rows(e.neighbor_pairs(r, 2.0))
    ref_pairs = _sorted_rows(o.nl_pairs(2.0))
    print(f"pairs: {len(gpu_pairs)} on the device, {len(ref_pairs)} in the oracle", flush=True)
    assert len(ref_pairs) > 3_000_000
    assert np.array_equal(gpu_pairs, ref_pairs)
    # reaction candidates (n_educts, handle1, handle2, reaction): bit-exact
    gpu_c = _sorted_rows(e.reaction_candidates(r))
    ref_c = _sorted_rows(o.reaction_candidates())
    print(f"candidates: {len(gpu_c)} / {len(ref_c)}", flush=True)
    assert len(ref_c) > 100_000 and np.array_equal(gpu_c, ref_c)
    # forces and energy within 1e-5 of FP64
    e.calculate_forces()
    f, fh = e.get_forces(r)
    e_ref = o.calculate_forces()
    _, _, ids, f_ref = o.particles()
    f_ref = f_ref[np.argsort(ids)]
    assert np.array_equal(fh, np.sort(ids).astype(np.uint32))
    ferr = np.linalg.norm(f - f_ref) / np.linalg.norm(f_ref)
    eerr = abs(e.energy(r) - e_ref) / abs(e_ref)
    print(f"force rel err {ferr:.2e}, energy rel err {eerr:.2e}", flush=True)
    assert ferr < 1e-5 and eerr < 1e-5
