"""Pins of the oracle's water-table bottom boundary of the Richards solve
(lower_boundary_condition = bc_waterTable: SoilWaterMovementMod.F90:1376-
1377, 1767-1808, compute_qcharge :2016-2135).

  - water table inside the soil column: no flow across the bottom, so the
    solve is the zero-flux solve bit for bit, and qcharge is the recharge of
    compute_qcharge, checked against a numpy evaluation on steps of dtmin
    (one trial, which sees the initial water), including
    the jwt = 0 case with its binary32 1.e-3 and the +-10 mm per step limit;
  - water table far below the column: the flux to the table tends to free
    drainage, so the solve tends to the flux-boundary solve;
  - the water balance of the solve closes with the sub-steps' drainage;
  - zwt on an interface counts as inside (zwt <= zi); a column whose
    nbedrock is not nlevsoi is refused.
"""
import numpy as np
import pytest

from ctsm_amd import synthetic as S
from oracle import oracle as O

NCOL = 256
NS = 12
BC_FLUX, BC_ZERO_FLUX, BC_WATERTABLE = 1, 2, 3     # CTSM_BC_* of include/ctsm_gpu.h


def _cfg(bc):
    cfg = O.config()
    cfg.lower_boundary_condition = bc
    return cfg


@pytest.fixture(scope="module")
def base():
    st = S.soil_only(NCOL, 91, snow=True, bedrock=False)
    assert np.all(st.col["nbedrock"] == 20)
    return st


def _run(st, bc, dtime=1800.0):
    o = st.copy()
    rc, bad = O.soilwater(o, _cfg(bc), dtime)
    assert rc == 0, bad
    return o


def _zi(st, j):
    return st.col["zi"][j + NS]                      # zi [-nlevsno : nlevgrnd]


def test_water_table_inside_is_the_zero_flux_solve(base):
    st = base.copy()
    rng = np.random.default_rng(1)
    j = rng.integers(1, 20, NCOL)                     # table inside layer j+1 (jwt = j)
    zi = np.stack([_zi(st, k) for k in range(0, 21)])   # zi(0..20)
    frac = rng.uniform(0.05, 0.95, NCOL)
    st.col["zwt"][:] = zi[j, np.arange(NCOL)] + frac * (zi[j + 1, np.arange(NCOL)] - zi[j, np.arange(NCOL)])
    wt = _run(st, BC_WATERTABLE)
    zf = _run(st, BC_ZERO_FLUX)
    for k in ("h2osoi_liq", "smp_l", "hk_l", "qin", "qout", "num_substeps"):
        np.testing.assert_array_equal(wt.col[k], zf.col[k], err_msg=k)
    assert np.all(zf.col["qcharge"] == 0.0) and np.any(wt.col["qcharge"] != 0.0)


def _numpy_qcharge(st, o, jwt, dtime):
    """compute_qcharge on the initial water (columns with one sub-step)."""
    c = np.arange(NCOL)
    col = st.col
    zwt = col["zwt"]
    l1 = jwt                                           # 0-based row of layer jwt+1 in SOI arrays
    ws = col["watsat"][l1, c]
    dz = col["dz"][NS + l1, c]
    liq = col["h2osoi_liq"][NS + l1, c]
    vwc = np.maximum(liq, 1.0e-6) / (dz * 1000.0)
    s1 = np.minimum(1.0, np.maximum(vwc / ws, 0.01))
    ic = col["icefrac"][l1, c]
    icn = col["icefrac"][np.minimum(l1 + 1, 19), c]
    imped = 10.0 ** (-(6.0 * (0.5 * (ic + icn))))
    ka = imped * col["hksat"][l1, c] * s1 ** (2.0 * col["bsw"][l1, c] + 3.0)
    jw = np.maximum(1, jwt)
    smp1 = np.maximum(col["smpmin"], o.col["smp_l"][jw - 1, c])
    wh_zwt = -col["sucsat"][l1, c] - zwt * 1000.0
    wh = smp1 - col["z"][NS + jw - 1, c] * 1000.0
    zj = col["z"][NS + np.maximum(jwt, 1) - 1, c]
    q = np.where(jwt == 0, -(ka * (wh_zwt - wh) / ((zwt + float(np.float32(1.e-3))) * 1000.0)),
                 -(ka * (wh_zwt - wh) / ((zwt - zj) * 1000.0 * 2.0)))
    return np.minimum(10.0 / dtime, np.maximum(-10.0 / dtime, q))


@pytest.mark.parametrize("where", ["top", "inside"])
def test_qcharge_is_compute_qcharge(base, where):
    st = base.copy()
    rng = np.random.default_rng(2 if where == "top" else 3)
    c = np.arange(NCOL)
    zi = np.stack([_zi(st, k) for k in range(0, 21)])
    if where == "top":
        jwt = np.zeros(NCOL, dtype=int)               # zwt above the first interface: jwt = 0
        st.col["zwt"][:] = zi[1, c] * rng.uniform(0.05, 0.95, NCOL)
    else:
        jwt = rng.integers(1, 20, NCOL)
        st.col["zwt"][:] = zi[jwt, c] + rng.uniform(0.05, 0.95, NCOL) * (zi[jwt + 1, c] - zi[jwt, c])
    # a step of dtmin (60 s) is one trial, accepted: its water is the initial water
    o = _run(st, BC_WATERTABLE, dtime=60.0)
    assert np.all(o.col["num_substeps"] == 1)
    ref = _numpy_qcharge(st, o, jwt, 60.0)
    np.testing.assert_allclose(o.col["qcharge"], ref, rtol=1e-12, atol=1e-300)
    assert np.all(np.abs(o.col["qcharge"]) <= 10.0 / 60.0) and np.any(o.col["qcharge"] != 0.0)


def test_qcharge_limit_binds():
    st = S.soil_only(64, 93, snow=False, bedrock=False)
    st.col["hksat"][...] *= 1e4
    st.col["zwt"][:] = 0.5 * st.col["zi"][NS + 1]       # just below the surface
    o = _run(st, BC_WATERTABLE)
    assert np.any(np.abs(o.col["qcharge"]) == 10.0 / 1800.0)


def test_deep_water_table_tends_to_free_drainage(base):
    st = base.copy()
    st.col["zwt"][:] = 1e9                             # m: the gradient term vanishes
    wt = _run(st, BC_WATERTABLE)
    fl = _run(st, BC_FLUX)
    np.testing.assert_allclose(wt.col["h2osoi_liq"], fl.col["h2osoi_liq"], rtol=1e-9, atol=1e-9)
    # (qcharge differs: the flux boundary's drainage takes dhkdw * dwat, the
    # water table's the linearised outflow qout + dqodw1 * dwat, :1365-1377)
    assert np.sum(wt.col["qcharge"] > 0.0) > 0.9 * NCOL      # drainage out of the column
    # a shallower table below the column pulls harder (suction toward the table is smaller)
    st2 = base.copy()
    st2.col["zwt"][:] = _zi(st2, 20) + 0.5
    near = _run(st2, BC_WATERTABLE)
    assert np.mean(near.col["qcharge"]) != np.mean(wt.col["qcharge"])


def test_bottom_flux_to_a_table_below(base):
    """Darcy flux from the lowest node (suction smp, depth z) to a saturated
    plane at zwt: q = hk * (1 - (0 - smp) / (zwt - z)), downward positive;
    the last trial's qout / hk_l / smp_l are the stored diagnostics."""
    st = base.copy()
    zwt = _zi(st, 20) + np.random.default_rng(6).uniform(0.01, 3.0, NCOL)
    st.col["zwt"][:] = zwt
    o = _run(st, BC_WATERTABLE)
    den = 1000.0 * (zwt - st.col["z"][NS + 19])
    hk, smp = o.col["hk_l"][19], o.col["smp_l"][19]
    np.testing.assert_allclose(o.col["qout"][19], hk * (1.0 + smp / den), rtol=1e-12, atol=1e-300)
    # a dry bottom layer over a near table draws water up (capillary rise), a wet one drains
    q = o.col["qout"][19]
    assert np.any(q < 0.0) and np.any(q > 0.0)
    assert np.all((q < 0.0) == (-smp > den))


def test_water_balance_closes_with_the_bottom_drainage(base):
    st = base.copy()
    st.col["zwt"][:] = _zi(st, 20) + np.random.default_rng(4).uniform(0.1, 5.0, NCOL)
    o = _run(st, BC_WATERTABLE)
    dt = 1800.0
    liq0 = st.col["h2osoi_liq"][NS:NS + 20].sum(axis=0)
    liq1 = o.col["h2osoi_liq"][NS:NS + 20].sum(axis=0)
    net = (st.col["qflx_infl"] - st.col["qflx_rootsoi"][:20].sum(axis=0) - o.col["qcharge"]) * dt
    # to the precision of the sub-steps' tridiagonal solves (the zero-flux
    # solve of the same columns leaves the same residual, < 1e-7 of the water)
    assert np.all(np.abs(liq1 - liq0 - net) <= 2e-7 * liq0)
    assert np.max(np.abs(o.col["qcharge"] * dt) / liq0) > 1e-4


def test_table_on_an_interface_is_inside(base):
    st = base.copy()
    st.col["zwt"][:] = _zi(st, 20)                     # zwt == zi(nlevsoi): jwt = nlevsoi - 1
    on = _run(st, BC_WATERTABLE)
    zf = _run(st, BC_ZERO_FLUX)
    np.testing.assert_array_equal(on.col["h2osoi_liq"], zf.col["h2osoi_liq"])
    st.col["zwt"][:] = np.nextafter(_zi(st, 20), np.inf)   # just below: flux to the table
    below = _run(st, BC_WATERTABLE)
    assert np.any(below.col["h2osoi_liq"] != zf.col["h2osoi_liq"])


def test_bedrock_column_refused():
    st = S.soil_only(16, 95, snow=False, bedrock=True)
    assert np.any(st.col["nbedrock"] < 20)
    rc, bad = O.soilwater(st.copy(), _cfg(BC_WATERTABLE))
    assert rc != 0 and st.col["nbedrock"][bad - 1] < 20
