"""Host logic of the ensemble server (no GPU): parameter perturbation bounds
(30 parameters, soil / surface columns and the member's own PFT rows, the
relations between them kept), the oracle stepping a perturbed member to
finite state, and request stream statistics."""
import math

import numpy as np

from ctsm_amd import serve, synthetic
from oracle import oracle as O


def test_perturbation_stays_physical():
    st = synthetic.with_canopy(synthetic.soil_only(64, 11), 11)
    base = synthetic.pft_table()
    ref = st.copy()
    f, pft = serve.perturb(st, 123)
    assert len(f) == 30 and len(serve.PERTURBED) == 30 and np.all((f >= 0.5) & (f <= 1.5))
    # parameters only: forcings and state other than the liquid clamp are untouched
    for k in ("forc_lwrad", "t_soisno", "h2osoi_ice", "qflx_infl", "t_grnd"):
        assert np.array_equal(st.col[k], ref.col[k], equal_nan=True), k
    for k in ("forc_u", "forc_v", "sabv", "parsun_z", "parsha_z", "dayl", "liqcan", "elai", "htop", "rootfr"):
        assert np.array_equal(st.can[k], ref.can[k], equal_nan=True), k
    c = st.col
    assert np.all(c["watsat"] <= 0.9) and np.all(c["emg"] <= 1.0)
    NS = st.dims.nlevsno
    assert np.all(c["h2osoi_liq"][NS:] <= c["watsat"] * c["dz"][NS:] * 1000.0)
    # the member's PFT table: every perturbed entry scaled by its factor, the
    # rest as the reference's rows; closure below opening
    for (grp, name, lo, hi), s in zip(serve.PERTURBED, f):
        if grp == "pft" and name not in ("smpsc", "fbw", "z0v_c"):
            assert np.allclose(pft[name], base[name] * s), name
    for name in ("c3psn", "theta_cj", "is_tree", "is_shrub", "fnitr"):
        assert np.array_equal(pft[name], base[name]), name
    assert np.all(pft["smpsc"] <= pft["smpso"] - 1000.0) and np.all(pft["fbw"] < 1.0)
    assert base["medlynslope"][3] == synthetic.pft_table()["medlynslope"][3]     # the base table is not modified


def test_perturbed_roughness_keeps_its_fixed_point():
    """z0v_c is held where the Meier (2022) fixed point U/u* = a exp(U/u*)
    has a root at every leaf+stem area up to z0v_LAImax, over many members;
    and a member whose unbounded z0v_c factor (1.45) took a = 0.45 > 1/e
    steps to finite canopy state in the oracle."""
    st = synthetic.with_canopy(synthetic.soil_only(4, 5), 5)
    for seed in range(400):
        _, pft = serve.perturb(st.copy(), seed)
        lt = pft["z0v_LAImax"]
        a = pft["z0v_c"] * lt / (4.0 * np.sqrt(pft["z0v_Cs"] + 0.5 * pft["z0v_Cr"] * lt))
        assert np.all(a <= 1.0 / np.e), seed
    q = serve.make_requests(6, rate=1e6, seed=21, max_cols=300, max_days=3, steps_per_day=2)[5]
    assert np.random.default_rng(q.seed).uniform(0.5, 1.5, 30)[23] > 1.4      # the z0v_c factor
    serve.build_member(q)
    o = q.state
    rc, bad = O.land_step_full(o, O.canopy_config(), synthetic.canopy_params(), O.PftStruct(q.pft),
                               hydro=O.hydro_params(), water_table_after=True, drainage_after=True)
    assert rc == 0, bad
    for k in ("t_veg", "rb1", "ram1", "canopy_err"):
        assert np.all(np.isfinite(o.can[k])), k


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
