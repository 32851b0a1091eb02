"""Serving loop: continuous batching of simulation requests on one device-resident engine.

Requests (a particle count, a step count, a reaction handler, a seed) arrive over time.  The
server admits a waiting request as soon as a replica slot, cells of the engine's pool and the HBM
budget allow, uploads its initial state once, and from then on it steps in the shared launches of
every admitted request (one hipGraph-replayed batch per scheduling interval), each replica with its
own step budget and handler.  Between intervals the server samples observables for all replicas at
once (particles per type, reaction totals), retires finished requests (final state read back, slot
freed) and admits new ones.  Observables go to HDF5 files in ReaDDy's layout on a writer thread,
off the critical path; a server built with record_trajectory also reads every replica's particles
back at each interval and writes them as a FlatTrajectory (one frame per interval), and one built
with record_reactions writes the reaction records of each interval's last step (the Reactions
observable with the interval as its stride).

Every request's trajectory depends only on its own seed and steps (replica-local handles and
Philox counters, fixed-point force sums), so a result is the same alone and in any mix.

The workload is the served-stream configuration: the 1M-particle system's chemistry (harmonic
repulsion between all species, A-A weak interaction, A+B<->C) at density 0.125, particle counts
log-uniform in [1000, 200000], step counts log-uniform in [1000, 100000] (scaled by `step_scale`
for finite traces), handlers 50/50, Poisson arrivals.
"""
import math
import os
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from . import h5io
from .engine import Engine

DENSITY = 0.125
DT = 0.005
DIFFUSION = [1.0, 1.0, 0.5]
HANDLERS = {"UncontrolledApproximation": L.HANDLER_UA, "Gillespie": L.HANDLER_GILLESPIE}


def chemistry():
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    pots.append((L.POT_WEAK_INTERACTION, 0, 0, [10.0, 1.0, 0.5, 2.0]))
    rxs = [dict(kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=0.1, educt_distance=1.0),
           dict(kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=0.05, product_distance=1.0)]
    return pots, rxs


@dataclass
class Request:
    rid: int
    n: int
    steps: int
    handler: str
    seed: int
    arrival: float  # seconds after the trace starts

    @property
    def box(self):
        return (self.n / DENSITY) ** (1.0 / 3.0)

    def initial_state(self):
        """Uniform positions, 40 % A / 40 % B / 20 % C, reproducible from the seed."""
        rng = np.random.default_rng([self.seed, self.rid])
        pos = (rng.random((self.n, 3)) - 0.5) * self.box
        types = rng.choice(np.array([0, 1, 2], dtype=np.uint32), self.n, p=[0.4, 0.4, 0.2])
        return pos, types


@dataclass
class Result:
    request: Request
    admitted: float = 0.0
    completed: float = 0.0
    particle_steps: float = 0.0
    final: tuple = None           # (positions, types, handles) sorted by handle
    observed_steps: list = field(default_factory=list)
    observed_counts: list = field(default_factory=list)
    observed_reactions: list = field(default_factory=list)
    observed_frames: list = field(default_factory=list)  # (handles, types, positions) per interval
    observed_records: list = field(default_factory=list)  # reaction records of each interval's last step

    @property
    def latency(self):
        return self.completed - self.request.arrival


def poisson_trace(n_requests, rate_per_s=200.0, n_range=(1000, 200000), steps_range=(1000, 100000),
                  step_scale=1.0, seed=4):
    """Poisson arrivals; N and steps log-uniform; handler 50/50."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate_per_s, n_requests))
    n = np.round(np.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1]), n_requests))).astype(int)
    steps = np.exp(rng.uniform(math.log(steps_range[0]), math.log(steps_range[1]), n_requests))
    steps = np.maximum(1, np.round(steps * step_scale)).astype(int)
    handler = np.where(rng.random(n_requests) < 0.5, "UncontrolledApproximation", "Gillespie")
    return [Request(i, int(n[i]), int(steps[i]), str(handler[i]), 1000 + i, float(t[i])) for i in range(n_requests)]


class H5Writer:
    """Writes each finished request's observables (and its trajectory, when recorded) to
    <dir>/request_<rid>.h5 on a background thread."""

    def __init__(self, directory):
        self.dir = directory
        os.makedirs(directory, exist_ok=True)
        h5io.lib()  # fails here, not on the writer thread, when the library is missing
        self.q = queue.Queue()
        self.errors = []
        self.written = []
        self.t = threading.Thread(target=self._run, daemon=True)
        self.t.start()

    def submit(self, result):
        self.q.put(result)

    def close(self):
        self.q.put(None)
        self.t.join()

    def _run(self):
        while True:
            res = self.q.get()
            if res is None:
                return
            try:
                path = os.path.join(self.dir, f"request_{res.request.rid}.h5")
                with h5io.File(path) as f:
                    f.config([(name, d) for name, d in zip("ABC", DIFFUSION)], [res.request.box] * 3,
                             reactions=[dict(rx, id=k, name=("fusion", "fission")[k])
                                        for k, rx in enumerate(chemistry()[1])])
                    t = res.observed_steps
                    f.n_particles("n_particles", t, res.observed_counts)
                    # totals since admission, in ReactionCounts' layout (the reference's observable holds
                    # the last step's counts, hence another name)
                    f.reaction_counts("reaction_totals", t, res.observed_reactions)
                    if res.observed_records:
                        f.reactions("reactions", t, res.observed_records)
                    if res.observed_frames:
                        # every particle of these requests is a normal one (flavor 0)
                        f.flat_trajectory("trajectory", t, [(h, ty, np.zeros(len(h), np.uint8), x)
                                                            for x, ty, h in res.observed_frames])
                    for k, v in (("handler", res.request.handler), ("particles", res.request.n),
                                 ("steps", res.request.steps), ("seed", res.request.seed), ("timestep", repr(DT))):
                        f.attribute(k, v)
                self.written.append(path)
            except Exception as exc:  # reported by the server after the run
                self.errors.append(exc)


class Server:
    """Continuous batching over one engine.

    hbm_budget_bytes bounds the sum of admitted requests' footprints (rdy_replica_footprint); the
    engine's cell pool is sized to the same budget, so an admitted request always finds its cells
    unless the pool is fragmented (then it waits for the next retirement).
    """

    def __init__(self, hbm_budget_bytes, max_slots=512, interval_steps=100, nl_cutoff=2.0, out_dir=None,
                 max_particles_per_request=200000, record_trajectory=False, record_reactions=False):
        pots, rxs = chemistry()
        self.interval = int(interval_steps)
        self.budget = int(hbm_budget_bytes)
        lam = DENSITY * nl_cutoff ** 3
        cap = max(8, int(math.ceil(lam + 7.0 * math.sqrt(lam) + 6.0)) // 4 * 4 + 4)
        bytes_per_cell = 2 * cap * 32 + 8
        max_cells = max(1, self.budget // bytes_per_cell)
        max_particles = int(max_cells * lam * 2) + 4096
        self.engine = Engine(3, DIFFUSION, potentials=pots, reactions=rxs, nl_cutoff=nl_cutoff, bin_capacity=cap,
                             max_replicas=max_slots, max_cells=max_cells, max_particles=max_particles)
        self.max_slots = max_slots
        self.writer = H5Writer(out_dir) if out_dir else None
        self.record_trajectory = bool(record_trajectory)
        self.record_reactions = bool(record_reactions)
        self.n_reactions = len(rxs)

    def footprint(self, req):
        return self.engine.footprint([req.box] * 3, 2 * req.n)

    def run(self, trace, keep_final=False, realtime=True):
        """Serves the trace; returns (results by request id, summary)."""
        eng = self.engine
        pending = sorted(trace, key=lambda r: r.arrival)
        waiting, active, results = [], {}, {}
        used = 0
        busy = 0.0
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        sim_time = 0.0
        while pending or waiting or active:
            now = clock() if realtime else sim_time
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))
            # admission, first come first served, by HBM budget and free slots / cells
            still = []
            for req in waiting:
                fp = self.footprint(req)
                if len(active) >= self.max_slots or used + fp > self.budget:
                    still.append(req)
                    continue
                try:
                    rid = eng.add_replica([req.box] * 3, (1, 1, 1), seed=req.seed, max_particles=2 * req.n)
                except L.EngineError:
                    still.append(req)  # cell pool fragmented: wait for a retirement
                    continue
                pos, types = req.initial_state()
                eng.set_particles(rid, pos, types)
                eng.set_handler(rid, HANDLERS[req.handler])
                eng.set_steps(rid, req.steps)
                if self.record_reactions:
                    eng.record_reactions(True)
                used += fp
                res = Result(req, admitted=now)
                results[req.rid] = res
                active[rid] = [req, 0, fp, res, req.n]
            waiting = still
            if not active:
                if pending:
                    if realtime:
                        time.sleep(max(0.0, pending[0].arrival - clock()))
                    else:
                        sim_time = pending[0].arrival
                    continue
                if waiting:
                    raise RuntimeError("a request does not fit the HBM budget even on an idle server")
                break
            # one batched interval: every admitted request advances up to its remaining steps
            k = min(self.interval, max(a[0].steps - a[1] for a in active.values()))
            tb = time.perf_counter()
            eng.run(L.HANDLER_UA, DT, k, graph=True)
            busy += time.perf_counter() - tb
            if not realtime:
                sim_time += time.perf_counter() - tb
            counts = eng.count_types(max(active) + 1)
            now = clock() if realtime else sim_time
            for r in list(active):
                req, done, fp, res, n_before = active[r]
                adv = min(k, req.steps - done)
                done += adv
                n_after = int(counts[r].sum())
                active[r][1] = done
                active[r][4] = n_after
                res.particle_steps += 0.5 * (n_before + n_after) * adv
                res.observed_steps.append(done)
                res.observed_counts.append(counts[r].copy())
                res.observed_reactions.append(eng.reaction_totals(r).copy())
                if self.record_trajectory:
                    res.observed_frames.append(eng.get_particles(r))
                if self.record_reactions:
                    res.observed_records.append(eng.reaction_records(r))
                if done >= req.steps:
                    if eng.steps_done(r) != req.steps:
                        raise RuntimeError(f"request {req.rid}: {eng.steps_done(r)} steps done, {req.steps} asked")
                    if keep_final:
                        res.final = eng.get_particles(r)
                    res.completed = now
                    eng.remove_replica(r)
                    used -= fp
                    del active[r]
                    if self.writer:
                        self.writer.submit(res)
        wall = (clock() if realtime else sim_time)
        if self.writer:
            self.writer.close()
            if self.writer.errors:
                raise self.writer.errors[0]
        lat = np.array([r.latency for r in results.values()])
        ps = sum(r.particle_steps for r in results.values())
        summary = {"requests": len(results), "wall_s": wall, "particle_steps": ps,
                   "particle_steps_per_s": ps / wall if wall > 0 else 0.0,
                   "busy_particle_steps_per_s": ps / busy if busy > 0 else 0.0,
                   "utilisation": busy / wall if wall > 0 else 0.0,
                   "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99))}
        return results, summary

    def close(self):
        self.engine.close()


def run_alone(req, nl_cutoff=2.0):
    """The same request on an engine of its own (for alone-versus-in-the-mix checks)."""
    pots, rxs = chemistry()
    e = Engine(3, DIFFUSION, potentials=pots, reactions=rxs, nl_cutoff=nl_cutoff)
    r = e.add_replica([req.box] * 3, (1, 1, 1), seed=req.seed, max_particles=2 * req.n)
    pos, types = req.initial_state()
    e.set_particles(r, pos, types)
    e.run(HANDLERS[req.handler], DT, req.steps)
    out = e.get_particles(r), e.reaction_totals(r)
    e.close()
    return out
