"""Host logic of the ensemble server (no GPU): slot allocator, parameter
perturbation bounds, request stream statistics."""
import math

import numpy as np

from ctsm_amd import serve, synthetic


def test_range_alloc_first_fit_and_merge():
    a = serve.RangeAlloc(100)
    s1, s2, s3 = a.alloc(30), a.alloc(30), a.alloc(30)
    assert (s1, s2, s3) == (0, 30, 60)
    assert a.alloc(20) == -1
    a.release(s2, 30)
    assert a.alloc(20) == 30
    a.release(30, 20)
    a.release(s1, 30)
    a.release(s3, 30)
    assert a.free == [(0, 100)]


def test_perturbation_stays_physical():
    st = synthetic.with_canopy(synthetic.soil_only(64, 11), 11)
    f = serve.perturb(st, 123)
    assert len(f) == 30 and np.all((f >= 0.5) & (f <= 1.5))
    c, g = st.col, st.can
    assert np.all(c["watsat"] <= 0.9) and np.all(c["emg"] <= 1.0)
    assert np.all(g["laisun"] + g["laisha"] <= g["elai"] * (1 + 1e-12))
    assert np.allclose(g["rootfr"].sum(axis=0), 1.0)
    NS = st.dims.nlevsno
    assert np.all(c["h2osoi_liq"][NS:] <= c["watsat"] * c["dz"][NS:] * 1000.0)


def test_request_stream_distributions():
    reqs = serve.make_requests(4000, rate=10.0, seed=3)
    t = np.array([q.arrival for q in reqs])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 0.1) < 0.01
    n = np.array([q.ncol for q in reqs])
    assert n.min() >= 1 and n.max() <= 20000
    assert abs(np.log(n).mean() - math.log(20000) / 2) < 0.2      # log-uniform
    s = np.array([q.steps for q in reqs])
    assert s.min() >= 48 and s.max() <= 365 * 48
