"""Serving loop: trace generation and request state on the CPU; on the GPU, continuous batching
gives every request the result it has alone, and observables land in HDF5 in ReaDDy's layout."""
import json
import os

import numpy as np
import pytest

from readdy_amd import h5io
from readdy_amd import serving as S



def test_trace_is_poisson_and_log_uniform():
    tr = S.poisson_trace(4000, rate_per_s=200.0)
    gaps = np.diff([r.arrival for r in tr])
    assert abs(gaps.mean() - 1 / 200.0) < 0.1 / 200.0       # Poisson(200/s): mean gap 5 ms
    ln = np.log([r.n for r in tr])
    assert ln.min() >= np.log(1000) - 1e-9 and ln.max() <= np.log(200000) + 1e-9
    assert abs(ln.mean() - 0.5 * (np.log(1000) + np.log(200000))) < 0.1
    frac_g = np.mean([r.handler == "Gillespie" for r in tr])
    assert 0.45 < frac_g < 0.55


def test_request_state_is_reproducible_and_in_the_box():
    r = S.Request(7, 5000, 10, "Gillespie", 123, 0.0)
    p1, t1 = r.initial_state()
    p2, t2 = r.initial_state()
    assert np.array_equal(p1, p2) and np.array_equal(t1, t2)
    assert np.all(np.abs(p1) <= r.box / 2)
    assert abs(r.n / r.box ** 3 - S.DENSITY) < 1e-12


def _h5_read(path, dataset):
    return h5io.read_u64(path, dataset)


@pytest.mark.gpu
def test_served_results_equal_results_alone(tmp_path):
    trace = S.poisson_trace(8, rate_per_s=400.0, n_range=(500, 4000), steps_range=(40, 400), seed=9)
    srv = S.Server(hbm_budget_bytes=256 << 20, max_slots=4, interval_steps=25, out_dir=str(tmp_path))
    results, summary = srv.run(trace, keep_final=True)
    srv.close()
    assert summary["requests"] == len(trace)
    for req in trace:
        res = results[req.rid]
        (xa, ta, ha), totals = S.run_alone(req)
        xm, tm, hm = res.final
        assert np.array_equal(ha, hm) and np.array_equal(ta, tm), req
        assert np.array_equal(xa, xm), req
        assert np.array_equal(totals, res.observed_reactions[-1]), req
        path = os.path.join(str(tmp_path), f"request_{req.rid}.h5")
        t = _h5_read(path, "/readdy/observables/n_particles/time")
        d = _h5_read(path, "/readdy/observables/n_particles/data").reshape(len(t), 3)
        assert t[-1] == req.steps and np.array_equal(d[-1], np.bincount(tm, minlength=3))
        c0 = _h5_read(path, "/readdy/observables/reaction_totals/counts/0")
        assert len(c0) == len(t) and c0[-1] == totals[0]
    # a budget that holds at most a few requests at a time made them queue and overlap
    assert summary["latency_p99_s"] >= summary["latency_p50_s"] > 0


@pytest.mark.gpu
def test_served_trajectory_frames(tmp_path):
    """With record_trajectory, request_<rid>.h5 holds one FlatTrajectory frame per observed interval:
    the frame times are the observed step counts, each frame's size is the particle count observed
    then, and the last frame is the request's final state (handles as ids, flavor 0).  With
    record_reactions, the Reactions observable holds the records of each interval's last step."""
    trace = S.poisson_trace(3, rate_per_s=400.0, n_range=(3000, 6000), steps_range=(30, 120), seed=12)
    srv = S.Server(hbm_budget_bytes=256 << 20, max_slots=2, interval_steps=5, out_dir=str(tmp_path),
                   record_trajectory=True, record_reactions=True)
    results, _ = srv.run(trace, keep_final=True)
    srv.close()
    n_records = 0
    for req in trace:
        res = results[req.rid]
        path = os.path.join(str(tmp_path), f"request_{req.rid}.h5")
        print("trajectory of request", req.rid, req.n, req.steps, flush=True)
        t, frames = h5io.read_trajectory(path, "trajectory")
        assert np.array_equal(t, res.observed_steps) and t[-1] == req.steps
        for (ids, ty, fl, x), c in zip(frames, res.observed_counts):
            assert len(ids) == int(np.sum(c)) and np.array_equal(np.bincount(ty, minlength=3), c)
            assert not fl.any()
        xf, tf, hf = res.final
        ids, ty, _, x = frames[-1]
        # the records of each interval's last step against the frame read right after that step:
        # fusion A+B->C and fission C->A+B products are live with their types, the educts are gone
        tr, recs = h5io.read_reactions(path, "reactions")
        assert np.array_equal(tr, t)
        n_records += sum(len(f) for f in recs)
        for (fids, fty, _, _), rec in zip(frames, recs):
            live = dict(zip(fids.tolist(), fty.tolist()))
            for one in rec:
                ed, pr = [int(v) for v in one["educts"]], [int(v) for v in one["products"]]
                assert all(v not in live for v in ed)
                if one["type"] == 1:
                    assert one["reaction"] == 0 and live[pr[0]] == 2 and sorted(one["types_from"]) == [0, 1]
                else:
                    assert one["type"] == 2 and one["reaction"] == 1 and list(one["types_from"]) == [2, 2]
                    assert live[pr[0]] == 0 and live[pr[1]] == 1
        general = json.loads(h5io.read_string(path, "/readdy/config/general"))
        assert general["box_size"] == [req.box] * 3 and general["pbc"] == [True] * 3
        assert np.array_equal(ids, hf) and np.array_equal(ty, tf) and np.array_equal(x, xf)
    assert n_records > 5  # about 0.5 events per step at these sizes, sampled every fifth step


def test_writer_thread_files(tmp_path):
    """The writer thread's file for a finished request (no GPU: a Result filled by hand)."""
    req = S.Request(3, 40, 60, "Gillespie", 77, 0.0)
    res = S.Result(req)
    rng = np.random.default_rng(0)
    for step, n in ((20, 40), (40, 39), (60, 41)):
        ty = rng.integers(0, 3, n).astype(np.uint32)
        res.observed_steps.append(step)
        res.observed_counts.append(np.bincount(ty, minlength=3))
        res.observed_reactions.append(np.array([step // 20, step // 30], dtype=np.uint64))
        res.observed_frames.append((rng.normal(size=(n, 3)), ty, np.arange(n, dtype=np.uint64) + step))
    w = S.H5Writer(str(tmp_path))
    w.submit(res)
    w.close()
    assert not w.errors, w.errors
    path = w.written[0]
    assert path.endswith("request_3.h5")
    t, frames = h5io.read_trajectory(path, "trajectory")
    assert list(t) == [20, 40, 60]
    for (x, ty, h), (ids, ty2, fl, x2) in zip(res.observed_frames, frames):
        assert np.array_equal(ids, h) and np.array_equal(ty2, ty) and np.array_equal(x2, x) and not fl.any()
    assert np.array_equal(h5io.read_u64(path, "/readdy/observables/n_particles/data").reshape(3, 3),
                          np.array(res.observed_counts))
    assert list(h5io.read_u64(path, "/readdy/observables/reaction_totals/counts/1")) == [0, 1, 2]
    assert json.loads(h5io.read_string(path, "/readdy/config/general"))["box_size"] == [req.box] * 3
