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


def c1_system(n=1_000_000, L=200.0, seed=1):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) - 0.5) * L
    types = rng.choice(np.array([0, 1, 2], dtype=np.uint32), n, p=[0.4, 0.4, 0.2])
    return pos, types


def c1_chemistry(L):
    pots = [(L.POT_HARMONIC_REPULSION, a, b, [10.0, 1.0]) for a in range(3) for b in range(a, 3)]
    pots.append((L.POT_WEAK_INTERACTION, 0, 0, [10.0, 1.0, 0.5, 2.0]))
    rxs = [dict(kind=L.RX_FUSION, educts=[0, 1], products=[2], rate=0.1, educt_distance=1.0),
           dict(kind=L.RX_FISSION, educts=[2], products=[0, 1], rate=0.05, product_distance=1.0)]
    return pots, rxs


def traffic_from_profiles():
    """HBM bytes per launch of the walk kernel from the committed rocprofv3 --pmc summary, if any."""
    p = os.path.join(ROOT, "profiles", PMC_PROFILE)
    if not os.path.exists(p):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(p).read())
    return float(m.group(1)) if m else None


def cpu_baseline(n_steps=5):
    """FP64 oracle on the same 1M-particle system, a few steps (bounded sample), on the host cores
    this process may use (at most 16: forces, integration and the candidate search split over
    OpenMP threads, as the reference's CPU kernel splits them over its thread pool)."""
    from oracle import pyoracle as po
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
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
    o.nl_setup(o.max_cutoff())
    o.calculate_forces()
    t0 = time.perf_counter()
    ps = 0
    for t in range(n_steps):
        ps += o.n_active
        o.loop_step(0.005, po.UA_KEYED, t)
    dt = time.perf_counter() - t0
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), model)
    except OSError:
        pass
    phases = {k: v / n_steps * 1e3 for k, v in o.phase_times().items()}
    return {"value": ps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n_steps} full steps of the 1M-particle workload (integrate, neighbour list, UA reactions, "
                      f"neighbour list, forces) by the FP64 C oracle on {threads} OpenMP threads; {dt:.1f} s",
            "phase_ms_per_step": phases, "cpu_model": model, "nproc": os.cpu_count()}


def ensemble_bench(n_replicas, steps):
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
    ids = list(range(0, n_replicas, max(1, n_replicas // 32)))
    p = E.Plain(n, box, nl_cutoff=best["nl_cutoff"], replica_ids=ids)
    plain = p.timed(3)
    p.close()
    return {"replicas": int(n_replicas), "particles": particles,
            "batched_particle_steps_per_s": rate, "plain_particle_steps_per_s": plain, "speedup": rate / plain,
            "step_ms_p50": lat["p50_ms"], "step_ms_p99": lat["p99_ms"],
            "batched_roofline_frac": rate * B_ALG / (HBM_PEAK_GBS * 1e9),
            "steps": steps, "autotune": {"chosen": best, "tried": tried},
            "plain_sample": f"{len(ids)} replicas evenly spaced through the sample, one engine each, 3 steps",
            "latency": "per-step wall time of the batched loop over chunks of 10 graph-replayed steps"}


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
    eng = Engine(3, [1.0, 1.0, 0.5], potentials=pots, reactions=rxs)
    r = eng.add_replica([200.0] * 3, (1, 1, 1), seed=1, max_particles=len(types))
    eng.set_particles(r, pos, types)
    dt = 0.005
    handler = L.HANDLER_UA
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
