"""Batched replica ensembles: many independent ReaDDy systems of uneven size stepped together.

The workload is the batched-ensemble configuration: replicas of the A+B<->C system (harmonic
repulsion k=10 d=1 between all species, fusion A+B->C rate 0.1 radius 1, fission C->A+B rate 0.05
distance 1, D_A=D_B=1, D_C=0.5, dt=0.01, UncontrolledApproximation) with particle counts drawn
log-uniform in [500, 50000] and box edges scaled to density 0.25, each replica with its own Philox
key.  `batched` packs every replica into one engine, so each step is one set of launches over all
of them (captured in a hipGraph); `plain` runs the same kernels once per replica per step (one
engine per replica), the path batching is measured against.

The autotuner picks the neighbour-list cell size and the brick shape for an ensemble by timing short
runs of each candidate (the per-cell bin capacity follows from the expected occupancy).
"""
import math
import time

import numpy as np

from . import _lib as L
from .engine import Engine

DENSITY = 0.25
DT = 0.01
DIFFUSION = [1.0, 1.0, 0.5]


def chemistry():
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    rxs = [dict(kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=0.1, educt_distance=1.0),
           dict(kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=0.05, product_distance=1.0)]
    return pots, rxs


def sample(n_replicas=4096, seed=3, n_min=500, n_max=50000):
    """Particle counts log-uniform in [n_min, n_max] and box edges at DENSITY."""
    rng = np.random.default_rng(seed)
    n = np.round(np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_replicas))).astype(np.int64)
    box = (n / DENSITY) ** (1.0 / 3.0)
    return n, box


def initial_state(n, box, replica, seed=3):
    """Uniform positions, half A and half B, reproducible per replica."""
    rng = np.random.default_rng([seed, replica])
    pos = (rng.random((int(n), 3)) - 0.5) * box
    types = np.zeros(int(n), dtype=np.uint32)
    types[int(n) // 2:] = 1
    return pos, types


def poisson_capacity(lam, n_cells, expected_overflows=1e-3):
    """Smallest bin capacity c with n_cells * P(Poisson(lam) > c) below expected_overflows; the
    overflow region takes the rare rest."""
    c, p, cdf = 0, math.exp(-lam), math.exp(-lam)
    while n_cells * max(0.0, 1.0 - cdf) > expected_overflows and c < 4096:
        c += 1
        p *= lam / c
        cdf += p
    return max(4, c)


def cells_of(box, cutoff):
    return int(max(1, math.floor(box / cutoff))) ** 3


class Batched:
    """Every replica of the sample in one engine."""

    def __init__(self, n, box, nl_cutoff=1.0, bin_capacity=0, seed=0x5EED, brick=(0, 0, 0), sample_seed=3):
        pots, rxs = chemistry()
        self.n = np.asarray(n)
        self.box = np.asarray(box)
        cells = sum(cells_of(b, nl_cutoff) for b in self.box)
        total = int(self.n.sum())
        lam = total / cells
        cap = bin_capacity or poisson_capacity(lam, cells)
        self.nl_cutoff, self.cap = nl_cutoff, cap
        self.engine = Engine(3, DIFFUSION, potentials=pots, reactions=rxs, nl_cutoff=nl_cutoff, bin_capacity=cap,
                             max_replicas=len(self.n), max_cells=cells, max_particles=2 * total + 4096, brick=brick)
        self.replicas = []
        for r, (nr, b) in enumerate(zip(self.n, self.box)):
            rid = self.engine.add_replica([b] * 3, (1, 1, 1), seed=(seed << 32) | r, max_particles=2 * int(nr))
            pos, types = initial_state(nr, b, r, sample_seed)
            self.engine.set_particles(rid, pos, types)
            self.replicas.append(rid)

    def particles(self):
        # one counting pass over all replicas (a gather per replica would read the whole pool 4096 times)
        return int(self.engine.count_types(max(self.replicas) + 1).sum())

    def run(self, steps, graph=True):
        self.engine.run(L.HANDLER_UA, DT, steps, graph=graph)

    def timed(self, steps, chunk=10, warmup=10):
        """(particle-steps/s, per-step wall-time percentiles over chunks of `chunk` steps).  The steps are one
        open run (its first walk and last copy once, as a serving loop keeps its run open), enqueued and
        synchronised chunk by chunk for the latencies."""
        self.run(warmup)
        n0 = self.particles()
        per_step = []
        e = self.engine
        t_all = time.perf_counter()
        e.run_begin(L.HANDLER_UA, DT, steps)
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            t0 = time.perf_counter()
            e.run_enqueue(k)
            e.synchronize()
            per_step.append((time.perf_counter() - t0) / k)
            done += k
        e.run_end()
        elapsed = time.perf_counter() - t_all
        n1 = self.particles()
        rate = 0.5 * (n0 + n1) * steps / elapsed
        ps = np.array(per_step)
        return rate, {"p50_ms": 1e3 * float(np.percentile(ps, 50)), "p99_ms": 1e3 * float(np.percentile(ps, 99)),
                      "mean_ms": 1e3 * float(ps.mean())}

    def close(self):
        self.engine.close()


class Plain:
    """The same replicas, one engine each, stepped one replica at a time (one launch sequence per
    replica per step)."""

    def __init__(self, n, box, nl_cutoff=1.0, seed=0x5EED, sample_seed=3, replica_ids=None):
        pots, rxs = chemistry()
        self.engines = []
        ids = range(len(n)) if replica_ids is None else replica_ids
        for r in ids:
            nr, b = int(n[r]), float(box[r])
            e = Engine(3, DIFFUSION, potentials=pots, reactions=rxs, nl_cutoff=nl_cutoff)
            rid = e.add_replica([b] * 3, (1, 1, 1), seed=(seed << 32) | r, max_particles=2 * nr)
            pos, types = initial_state(nr, b, r, sample_seed)
            e.set_particles(rid, pos, types)
            self.engines.append((e, rid))

    def particles(self):
        return sum(e.n_particles(r) for e, r in self.engines)

    def timed(self, steps, warmup=1):
        """Particle-steps/s of a run of `steps` per replica, each replica's step kernels launched once per
        step with the replicas interleaved on their engines' streams and one synchronisation at the end
        (the run's first walk and last copy included)."""
        for e, _ in self.engines:
            e.run(L.HANDLER_UA, DT, warmup)
        n0 = self.particles()
        t0 = time.perf_counter()
        for e, _ in self.engines:
            e.run_begin(L.HANDLER_UA, DT, steps)
        for _ in range(steps):
            for e, _ in self.engines:
                e.run_enqueue(1)
        for e, _ in self.engines:
            e.run_end()
        elapsed = time.perf_counter() - t0
        n1 = self.particles()
        return 0.5 * (n0 + n1) * steps / elapsed

    def timed_per_call(self, steps, warmup=1):
        """The earlier measure of the plain path: one whole run (first walk, step, last copy and a host
        synchronisation) per replica per step."""
        for e, _ in self.engines:
            e.run(L.HANDLER_UA, DT, warmup)
        n0 = self.particles()
        t0 = time.perf_counter()
        for _ in range(steps):
            for e, _ in self.engines:
                e.run(L.HANDLER_UA, DT, 1)
        elapsed = time.perf_counter() - t0
        n1 = self.particles()
        return 0.5 * (n0 + n1) * steps / elapsed

    def close(self):
        for e, _ in self.engines:
            e.close()


def half_round_brick(cell):
    """A cube of cells whose expected particles fill about half of a brick's 256 lanes (two waves): the
    engine's own choice fills one round of lanes, and with many small replicas the smaller brick can
    win (fewer waves per block, so a block's slowest wave holds its LDS for less time)."""
    lam = DENSITY * cell ** 3
    e = int(round((128.0 / lam) ** (1.0 / 3.0)))
    e = max(2, min(16, e))
    return (e, e, e)


def brick_candidates(cell):
    """The engine's one-round choice, the half-round cube and the two shapes next to it: how bricks tile
    the uneven replica grids (the last brick of an axis is clipped) changes the best shape by a cell."""
    e = half_round_brick(cell)[0]
    return [(0, 0, 0), (e, e, e), (e, e, e + 1), (e + 1, e + 1, e + 1)]


def autotune(n, box, cells=(1.0, 1.5, 2.0), steps=20):
    """Times short batched runs of each (cell size, brick shape) candidate (brick_candidates); returns the
    fastest with all timings."""
    results = []
    for cutoff in cells:
        for brick in brick_candidates(cutoff):
            b = Batched(n, box, nl_cutoff=cutoff, brick=brick)
            try:
                rate, _ = b.timed(steps, chunk=steps, warmup=5)
                results.append({"nl_cutoff": cutoff, "brick": list(brick), "bin_capacity": b.cap,
                                "particle_steps_per_s": rate})
            finally:
                b.close()
    best = max(results, key=lambda x: x["particle_steps_per_s"])
    return best, results
