"""Benchmark: particle-steps/s of the fused MI355X ReaDDy step on the 1M-particle A+B<->C system,
and the batched 4096-replica ensemble against the plain one-launch-per-replica path.

Workload (the headline configuration): periodic 200^3 box, 1,000,000 particles (40% A, 40% B,
20% C, uniform random, seed 1), harmonic repulsion k=10 d=1.0 between all species pairs,
weak-interaction piecewise harmonic A-A (k=10, d0=1.0, depth 0.5, cutoff 2.0), fusion A+B->C
rate 0.1 radius 1.0, fission C->A+B rate 0.05 distance 1.0, D_A = D_B = 1, D_C = 0.5,
EulerBD dt=0.005, UncontrolledApproximation.  FP32 pair arithmetic on FP64 positions.

Prints one JSON line.  `value` is whole-job particle-steps/s with the state resident in HBM;
`roofline` prices the step's dominant kernel, the stencil walk k_brick_step (pair forces, reaction
candidates and the next Euler-Maruyama move of one brick of cells), at 64 algorithmic bytes per
particle (its FP64 position, species and handle read once and written once), against MI355X's
8 TB/s; `step_roofline_frac` prices the whole step the same way.  `cpu_baseline` is the FP64 CPU oracle timed on a bounded sample of the same
workload on this host.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1M particles; batched ensemble throughput + p99 vs plain"
HBM_PEAK_GBS = 8000.0
B_ALG = 64.0  # bytes per particle-step: FP64 position + uint32 species + uint32 handle, read + written
B_WALK = 64.0  # bytes per particle of the stencil-walk kernel: the same record, read once and written once
PMC_PROFILE = "brick_step_c1_1M_pmc.md"
WALK_KERNEL = "k_brick_step<true> (pair forces, reaction candidates and the next move of one brick of cells)"


from readdy_amd.workloads import c1_chemistry, c1_system  # noqa: E402


def traffic_from_profiles():
    """HBM bytes per launch of the walk kernel from the committed rocprofv3 --pmc summary, if any."""
    p = os.path.join(ROOT, "profiles", PMC_PROFILE)
    if not os.path.exists(p):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(p).read())
    return float(m.group(1)) if m else None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            return next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "unknown")
    except OSError:
        return "unknown"


def _timed_loop(o, n_steps, dt, mode, step0=0):
    """(particle-steps/s, ms per phase and step) of n_steps full loop steps of oracle system o"""
    o.phase_times(reset=True)
    t0 = time.perf_counter()
    ps = 0
    for t in range(n_steps):
        ps += o.n_active
        o.loop_step(dt, mode, step0 + t)
    el = time.perf_counter() - t0
    return ps / el, el, {k: v / n_steps * 1e3 for k, v in o.phase_times().items()}


def _c1_oracle(threads):
    from oracle import pyoracle as po
    pos, types = c1_system()
    o = po.System(3, [200.0] * 3, [1, 1, 1], diffusion=[1.0, 1.0, 0.5])
    o.set_seed(1)
    for a in range(3):
        for b in range(a, 3):
            o.add_pair_potential(po.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0])
    o.add_pair_potential(po.POT_WEAK_INTERACTION, 0, 0, [10.0, 1.0, 0.5, 2.0])
    o.add_reaction(po.RX_FUSION, [0, 1], [2], 0.1, 1.0)
    o.add_reaction(po.RX_FISSION, [2], [0, 1], 0.05, 0.0, 1.0)
    o.add_particles(pos, types.astype(np.int32), ids=np.arange(len(types)))
    o.set_threads(threads)
    o.set_cas_fill(True)
    o.nl_setup(o.max_cutoff())
    o.calculate_forces()
    return o


def cpu_baseline(n_steps=5):
    """The FP64 oracle, this repository's restatement of the reference CPU kernel, timed on the host cores
    this process may use (at most 16): forces, integration and the candidate search split over OpenMP
    threads by static partitions, the cell list filled from all threads by atomic exchange, as the
    reference's CPU kernel does over its thread pool.  The headline figure is 5 full loop steps of the
    1M-particle workload on 1 x cores threads; beside it the reference's default pool of 4 x cores threads,
    the plain-method system in full (1000 steps), a sample of the bead-chain system and a sample of the
    batched ensemble run replica after replica."""
    from oracle import pyoracle as po
    from readdy_amd import ensemble as E
    from readdy_amd import workloads as W
    from readdy_amd import _lib as L
    cores = max(1, min(16, len(os.sched_getaffinity(0))))
    o = _c1_oracle(cores)
    rate, el, phases = _timed_loop(o, n_steps, 0.005, po.UA_KEYED)
    del o
    o = _c1_oracle(4 * cores)
    rate4, el4, _ = _timed_loop(o, n_steps, 0.005, po.UA_KEYED)
    del o
    # the plain-method system (2000 particles, box 20, UA, dt 0.01) in full: 1000 steps on one thread (on
    # this size more threads only add their overhead)
    rng = np.random.default_rng(0)
    c0 = po.System(3, [20.0] * 3, [1, 1, 1], diffusion=[1.0, 1.0, 0.5])
    for a in range(3):
        for b in range(a, 3):
            c0.add_pair_potential(po.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0])
    c0.add_reaction(po.RX_FUSION, [0, 1], [2], 0.1, 1.0)
    c0.add_reaction(po.RX_FISSION, [2], [0, 1], 0.05, 0.0, 1.0)
    c0.add_particles((rng.random((2000, 3)) - 0.5) * 20.0, np.repeat(np.arange(2, dtype=np.int32), 1000),
                     ids=np.arange(2000))
    c0.nl_setup(c0.max_cutoff())
    c0.calculate_forces()
    rate_c0, el_c0, ph_c0 = _timed_loop(c0, 1000, 0.01, po.UA_KEYED)
    # the bead-chain system: 10^4 chains of 50 beads + 5 10^5 crowders (box 160), 3 loop steps at
    # dt = 2e-4 (the configuration's 1e-3 blows the chains up), Gillespie
    pos, types, bonds, angles, dihedrals = W.system(10000, 50, 500000, 160.0, seed=2)
    c2 = po.System(3, [160.0] * 3, [1, 1, 1], diffusion=[0.5, 1.0, 1.0])
    for kind, a, b, p in W.potentials(L):
        c2.add_pair_potential(kind, a, b, p)
    for x in W.reactions(L, 0.01):
        c2.add_reaction(x["kind"], x["educts"], x["products"], x["rate"], x["educt_distance"], 0.0)
    c2.add_particles(pos, types.astype(np.int32), ids=np.arange(len(types)))
    for b in bonds:
        c2.add_bond(int(b[0]), int(b[1]), 100.0, 1.0)
    for a in angles:
        c2.add_angle(int(a[0]), int(a[1]), int(a[2]), 10.0, np.pi)
    for d in dihedrals:
        c2.add_dihedral(int(d[0]), int(d[1]), int(d[2]), int(d[3]), 1.0, 3, 0.0)
    c2.set_threads(cores)
    c2.set_cas_fill(True)
    c2.nl_setup(c2.max_cutoff())
    c2.calculate_forces()
    rate_c2, el_c2, _ = _timed_loop(c2, 3, 2e-4, po.GILLESPIE_KEYED)
    del c2
    # the batched ensemble, replica after replica: 8 replicas evenly spaced through the 4096, 100 steps each
    n, box = E.sample(4096)
    ps3, el3 = 0, 0.0
    for r in range(0, 4096, 512):
        o3 = po.System(3, [box[r]] * 3, [1, 1, 1], diffusion=E.DIFFUSION)
        for kind, a, b, p in E.chemistry()[0]:
            o3.add_pair_potential(kind, a, b, p)
        for x in E.chemistry()[1]:
            o3.add_reaction(x["kind"], x["educts"], x["products"], x["rate"], x.get("educt_distance", 0.0),
                            x.get("product_distance", 0.0))
        p3, t3 = E.initial_state(n[r], box[r], r)
        o3.add_particles(p3, t3.astype(np.int32), ids=np.arange(len(t3)))
        o3.nl_setup(o3.max_cutoff())
        o3.calculate_forces()
        rr, ee, _ = _timed_loop(o3, 100, E.DT, po.UA_KEYED)
        ps3 += rr * ee
        el3 += ee
    return {"value": rate, "unit": "particle-steps/s", "cores": cores, "kind": "port",
            "sample": f"{n_steps} full steps of the 1M-particle workload (integrate, neighbour list, UA reactions, "
                      f"neighbour list, forces) by the FP64 C oracle on {cores} OpenMP threads; {el:.1f} s",
            "phase_ms_per_step": phases, "cpu_model": _cpu_model(), "nproc": os.cpu_count(),
            "threads_4x_cores": {"threads": 4 * cores, "value": rate4, "seconds": el4},
            "plain_system_full": {"particles": 2000, "steps": 1000, "threads": 1, "value": rate_c0,
                                  "seconds": el_c0, "phase_ms_per_step": ph_c0},
            "bead_chain_sample": {"particles": int(len(types)), "steps": 3, "dt": 2e-4, "threads": cores,
                                  "value": rate_c2, "seconds": el_c2},
            "ensemble_sample": {"replicas": 8, "steps": 100, "threads": 1, "value": ps3 / el3, "seconds": el3}}


def ensemble_bench(n_replicas, steps, plain_steps=20):
    """Batched replica ensemble (4096 replicas of the A+B<->C system, 500..50000 particles each at
    density 0.25) against the plain one-engine-per-replica path; cell size and brick shape from the
    autotuner."""
    from readdy_amd import ensemble as E
    n, box = E.sample(n_replicas)
    sub = max(1, min(512, n_replicas))
    best, tried = E.autotune(n[:sub], box[:sub], steps=20)
    b = E.Batched(n, box, nl_cutoff=best["nl_cutoff"], brick=tuple(best["brick"]))
    rate, lat = b.timed(steps, chunk=10, warmup=10)
    particles = int(b.particles())
    b.close()
    # the plain path: one engine per replica, each replica's step kernels launched once per step on its
    # engine's stream, the replicas interleaved, one synchronisation per run (256 replicas evenly spaced
    # through the sample, 20 steps)
    ids = list(range(0, n_replicas, max(1, n_replicas // 256)))
    p = E.Plain(n, box, nl_cutoff=best["nl_cutoff"], replica_ids=ids)
    plain = p.timed(plain_steps)
    old_ids = ids[::max(1, len(ids) // 32)]
    p.engines = [p.engines[ids.index(i)] for i in old_ids]
    plain_per_call = p.timed_per_call(3)
    p.close()
    return {"replicas": int(n_replicas), "particles": particles,
            "batched_particle_steps_per_s": rate, "plain_particle_steps_per_s": plain, "speedup": rate / plain,
            "plain_per_call_particle_steps_per_s": plain_per_call, "speedup_vs_per_call": rate / plain_per_call,
            "step_ms_p50": lat["p50_ms"], "step_ms_p99": lat["p99_ms"],
            "batched_roofline_frac": rate * B_ALG / (HBM_PEAK_GBS * 1e9),
            "steps": steps, "autotune": {"chosen": best, "tried": tried},
            "plain_sample": f"{len(ids)} replicas evenly spaced through the sample, one engine each, a run of "
                            f"{plain_steps} steps enqueued step by step across the engines, one synchronisation",
            "plain_per_call_sample": f"{len(old_ids)} of them, 3 steps, one whole run (first walk, step, last copy, "
                                     f"host synchronisation) per replica per step",
            "sampled": f"{steps} of the configuration's 5000 steps",
            "latency": "per-step wall time of the batched loop: one open run of the sampled steps, enqueued and "
                       "synchronised in chunks of 10"}


def rdf_bench(eng, r, pos, types, sample=10000):
    """Radial distribution of the 1M-particle state (all species, 100 bins up to 5.0) on the device, and the
    reference's all-pairs host loop (RadialDistribution.cpp:75-103) as the oracle runs it on a sample of
    `sample` particles, scaled by (N / sample)^2 (the loop's cost depends on the pair count only)."""
    from oracle import pyoracle as po
    borders = np.linspace(0.0, 5.0, 101)
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        counts, nf = eng.rdf_counts(r, borders, [0, 1, 2], [0, 1, 2])
        times.append(time.perf_counter() - t0)
    dev = float(np.median(times))
    o = po.System(3, [200.0] * 3, [1, 1, 1])
    o.add_particles(pos[:sample], types[:sample].astype(np.int32), ids=np.arange(sample))
    t0 = time.perf_counter()
    o.rdf_counts(borders, [0, 1, 2], [0, 1, 2])
    host = (time.perf_counter() - t0) * (len(types) / sample) ** 2
    return {"particles": int(nf), "bins": len(borders) - 1, "r_max": 5.0, "device_s": dev,
            "pairs_counted": int(counts.sum()), "host_all_pairs_s_extrapolated": host,
            "speedup": host / dev,
            "host_sample": f"oracle all-pairs loop on {sample} particles, one core, scaled by (N/{sample})^2"}


def serving_bench(n_requests, step_scale):
    """Continuous batching of a finite Poisson trace (200 req/s) of the 1M-particle system's chemistry
    at 1000..200000 particles, steps 1000..100000 scaled by step_scale, handlers 50/50, 32 GiB HBM budget,
    observables written to HDF5 off the critical path."""
    import tempfile
    from readdy_amd import serving as S
    trace = S.poisson_trace(n_requests, step_scale=step_scale)
    with tempfile.TemporaryDirectory() as d:
        srv = S.Server(hbm_budget_bytes=32 << 30, max_slots=512, interval_steps=100, out_dir=d)
        _, summary = srv.run(trace)
        srv.close()
    summary.update({"trace": f"first {n_requests} arrivals of Poisson(200/s); steps log-uniform [1000, 100000] "
                             f"x {step_scale}", "hbm_budget_gib": 32,
                    "sampled": f"step counts scaled by {step_scale} so the trace finishes in seconds",
                    "roofline_frac": summary["particle_steps_per_s"] * B_ALG / (HBM_PEAK_GBS * 1e9)})
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--eager", action="store_true", help="launch kernels one by one instead of a hipGraph")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--profile-steps", type=int, default=200)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-rdf", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="use the engine's own brick for the 1M system")
    ap.add_argument("--serving-requests", type=int, default=200)
    ap.add_argument("--serving-step-scale", type=float, default=0.1)
    ap.add_argument("--ensemble-replicas", type=int, default=4096)
    ap.add_argument("--ensemble-steps", type=int, default=50)
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="first run one plain device copy of this many MiB (a known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build measures a single MI355X (--gpus {args.gpus} requested). "
              f"Replica ensembles shard across GPUs by replica; see the sharding notes in the design document.",
              file=sys.stderr)
        sys.exit(2)

    from readdy_amd import Engine, device_info
    from readdy_amd import _lib as L

    if args.calibrate_copy_mib:
        import torch
        n = args.calibrate_copy_mib * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    pos, types = c1_system()
    pots, rxs = c1_chemistry(L)
    dt = 0.005
    handler = L.HANDLER_UA

    def make_engine(cand):
        eng = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs, cell_radius=cand["cell_radius"],
                     brick=tuple(cand["brick"]))
        r = eng.add_replica([200.0] * 3, (1, 1, 1), seed=1, max_particles=len(types))
        eng.set_particles(r, pos, types)
        return eng, r

    tuned = None
    chosen = {"cell_radius": 1, "brick": [0, 0, 0]}
    if not args.no_autotune:
        from readdy_amd.tune import autotune_system, brick_candidates
        best, tried = autotune_system(make_engine, handler, dt, brick_candidates(len(types) / 200.0 ** 3, 2.0),
                                      steps=30, warmup=5)
        chosen = {"cell_radius": best["cell_radius"], "brick": best["brick"]}
        tuned = {"chosen": best, "tried": tried, "steps": 30}
        print(f"bench: autotuned C1 geometry {chosen}", file=sys.stderr, flush=True)
    eng, r = make_engine(chosen)
    graph = not args.eager
    eng.run(handler, dt, args.warmup, graph=graph)
    eng.synchronize()
    print(f"bench: warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    n0 = eng.n_particles(r)
    t0 = time.perf_counter()
    eng.run(handler, dt, args.steps, graph=graph)  # returns synchronized
    elapsed = time.perf_counter() - t0
    n1 = eng.n_particles(r)
    particle_steps = 0.5 * (n0 + n1) * args.steps
    value = particle_steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps

    # walk-kernel duration with HIP events on the engine's stream
    eng.profile_reset()
    eng.run(handler, dt, args.profile_steps, graph=False)
    avg_us, launches = eng.profile_get()
    n_mid = eng.n_particles(r)
    achieved = (n_mid * B_WALK) / (avg_us * 1e-6) / 1e9  # GB/s
    traffic = traffic_from_profiles()

    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp64 positions, fp32 pair forces",
        "data": "synthetic (uniform random positions, seed 1)",
        "config": {"workload": "A+B<->C, 1,000,000 particles in a periodic 200^3 box, UncontrolledApproximation, "
                               "dt=0.005, harmonic repulsion + A-A weak interaction", "particles": int(n0),
                   "box": 200.0, "handler": "UncontrolledApproximation", "graph": graph},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": WALK_KERNEL, "algorithmic_bytes_per_particle": B_WALK,
                     "avg_kernel_us": avg_us, "launches": launches},
        "step_roofline_frac": (value * B_ALG / 1e9) / HBM_PEAK_GBS,
        "geometry": {"cell_radius": chosen["cell_radius"], "brick": chosen["brick"], "autotune": tuned},
        "device": device_info()["name"],
    }
    if not args.no_rdf:
        out["rdf"] = rdf_bench(eng, r, pos, types)
        print(f"bench: radial distribution {out['rdf']['device_s'] * 1e3:.3g} ms", file=sys.stderr, flush=True)
    eng.close()
    print(f"bench: 1M-particle system {value:.4g} particle-steps/s", file=sys.stderr, flush=True)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(args.ensemble_replicas, args.ensemble_steps)
        print(f"bench: ensemble {out['ensemble']['batched_particle_steps_per_s']:.4g} particle-steps/s", file=sys.stderr,
              flush=True)
    if not args.no_serving:
        out["serving"] = serving_bench(args.serving_requests, args.serving_step_scale)
        print(f"bench: serving {out['serving']['particle_steps_per_s']:.4g} particle-steps/s", file=sys.stderr,
              flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
