"""Serving loop: trace generation and request state on the CPU; on the GPU, continuous batching
gives every request the result it has alone, and observables land in HDF5 in ReaDDy's layout."""
import ctypes
import os

import numpy as np
import pytest

from readdy_amd import serving as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


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
    lib = ctypes.CDLL(os.path.join(ROOT, "readdy_amd", "libreaddy_h5.so"))
    out = np.zeros(1 << 16, dtype=np.uint64)
    n = ctypes.c_uint64()
    rc = lib.rdy_h5_read_u64(path.encode(), dataset.encode(), ctypes.c_uint64(len(out)),
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ctypes.byref(n))
    assert rc == 0, (path, dataset)
    return out[:n.value]


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
        c0 = _h5_read(path, "/readdy/observables/reaction_counts/counts/0")
        assert len(c0) == len(t) and c0[-1] == totals[0]
    # a budget that holds at most a few requests at a time made them queue and overlap
    assert summary["latency_p99_s"] >= summary["latency_p50_s"] > 0
