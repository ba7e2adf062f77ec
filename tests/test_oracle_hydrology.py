"""Pins of the surface water oracle (oracle/hydrology_oracle.c) on the CPU.

The surface water input chain (HydrologyNoDrainageMod.F90:306-336) has no
golden vectors in the reference's tests; it is pinned by what the physics and
the routines' closed forms fix, each evaluated independently with numpy:
  * water balance of the surface: what reaches the ground (qflx_top_soil less
    the evaporation SetQflxInputs removes) = qflx_infl + qflx_surf + the
    change of h2osfc over the step, column by column, for both h2osfcflag
    values — a dropped or mis-routed flux breaks it;
  * h2osfc never negative after UpdateH2osfc, and with h2osfcflag = 0 the
    pool is emptied into the soil every step;
  * TOPMODEL fsat (both water-table branches), qinmax (minimum over layers
    1..3 with ice impedance), the SetQflxInputs partition (snow and snow-free
    branches), the overflow of the pool (QflxH2osfcSurf, including the
    per-step cap and the 1e-8 cut) and the drainage limit, against numpy;
  * special cases: no surface water fraction, frac_h2osfc_nosnow <= pc,
    nothing falling and nothing evaporating;
  * p2c of the three evaporation fluxes against a numpy weighted sum.
"""
import numpy as np
import pytest

from ctsm_amd import synthetic as S
from oracle import oracle as O

DT = 1800.0


def _state(ncol=512, seed=0x5EED0011, snow=True, fh=True, deep=True):
    """Soil columns with a surface water fraction and standing water that
    overflows its threshold in about half of the columns."""
    cols = np.arange(ncol, dtype=np.int64)
    frac = 0.6 * S.uniform(seed, 90, cols) if fh else None
    st = S.soil_only(ncol, seed, snow=snow, bedrock=True, frac_h2osfc=frac)
    S.with_hydrology(st, seed)
    if deep:
        st.col["h2osfc"][:] = 100.0 * S.uniform(seed, 91, cols)
    for j in range(3):      # some ice in the top layers, so the impedance term matters
        st.col["icefrac"][j] = np.where(S.uniform(seed, 92, cols, j) < 0.3, S.uniform(seed, 93, cols, j), 0.0)
    return st


def _hy(st):
    """All surface water fields; qflx_infl and h2osfc live in the column state."""
    return dict(st.hy, qflx_infl=st.col["qflx_infl"], h2osfc=st.col["h2osfc"])


def _run(st, hydro, dtime=DT):
    h0 = st.col["h2osfc"].copy()
    assert O.surface_water(st, hydro, dtime=dtime) == 0
    return h0


def _evap_removed(st):
    c, h = st.col, st.hy
    snow = c["snl"] < 0
    fsno = np.where(snow, c["frac_sno_eff"], 0.0)
    qevap = np.where(snow, h["qflx_ev_soil"], h["qflx_liqevap_from_top_layer"])
    fh = c["frac_h2osfc"]
    return (1.0 - fsno - fh) * qevap + fh * h["qflx_ev_h2osfc"]


def _fsat(st, fff):
    h = st.hy
    perched = (h["frost_table"] > h["zwt_perched"]) & (h["frost_table"] <= h["zwt"])
    z = np.where(perched, h["zwt_perched"], h["zwt"])
    return h["wtfact"] * np.exp(-0.5 * fff * z), perched


@pytest.mark.parametrize("flag", [1, 0])
def test_surface_water_balance(flag):
    st = _state()
    hydro = O.hydro_params(h2osfcflag=flag)
    h0 = _run(st, hydro)
    h = _hy(st)
    lhs = h["qflx_top_soil"] - _evap_removed(st)
    rhs = h["qflx_infl"] + h["qflx_surf"] + (st.col["h2osfc"] - h0) / DT
    scale = np.abs(h["qflx_top_soil"]) + np.abs(h0) / DT + 1e-12
    assert np.max(np.abs(lhs - rhs) / scale) < 1e-12
    assert np.all(st.col["h2osfc"] >= 0.0)
    # the routes out of the surface add up
    assert np.array_equal(h["qflx_infl"], h["qflx_in_soil_limited"] + h["qflx_h2osfc_drain"])
    assert np.array_equal(h["qflx_surf"],
                          h["qflx_sat_excess_surf"] + h["qflx_infl_excess_surf"] + h["qflx_h2osfc_surf"])


def test_no_surface_water_pool_drains_everything():
    st = _state()
    _run(st, O.hydro_params(h2osfcflag=0))
    h = _hy(st)
    assert np.all(st.col["h2osfc"] == 0.0)
    assert np.all(h["qflx_h2osfc_surf"] == 0.0)
    assert np.array_equal(h["qflx_infl_excess_surf"], h["qflx_infl_excess"])
    assert np.array_equal(h["qflx_in_h2osfc"], h["qflx_top_soil_to_h2osfc"])


def test_fsat_and_qinmax_closed_forms():
    st = _state()
    hydro = O.hydro_params(fff=0.5)
    _run(st, hydro)
    h, c = _hy(st), st.col
    fsat, perched = _fsat(st, 0.5)
    assert perched.any() and (~perched).any()
    assert np.allclose(h["fsat"], fsat, rtol=1e-15, atol=0)
    assert np.array_equal(h["fcov"], h["fsat"])
    assert np.allclose(h["qflx_sat_excess_surf"], fsat * h["qflx_rain_plus_snomelt"], rtol=1e-15, atol=0)
    e_ice = O.canopy_config().e_ice
    imp = 10.0 ** (-e_ice * c["icefrac"][:3]) * c["hksat"][:3]
    assert (c["icefrac"][:3] > 0).any()
    qinmax = (1.0 - fsat) * imp.min(axis=0)
    assert np.allclose(h["qinmax"], qinmax, rtol=1e-14, atol=0)
    excess = np.maximum(0.0, h["qflx_in_soil"] - (1.0 - c["frac_h2osfc"]) * qinmax)
    assert np.allclose(h["qflx_infl_excess"], excess, rtol=1e-12, atol=1e-20)
    assert (h["qflx_infl_excess"] > 0).any() and (h["qflx_infl_excess"] == 0).any()


def test_input_partition_closed_form():
    st = _state()
    _run(st, O.hydro_params())
    h, c = _hy(st), st.col
    snow = c["snl"] < 0
    assert snow.any() and (~snow).any()
    top = h["qflx_rain_plus_snomelt"] + h["qflx_snow_h2osfc"] + h["qflx_floodc"]
    assert np.array_equal(h["qflx_top_soil"], top)
    fh = c["frac_h2osfc"]
    fsno = np.where(snow, c["frac_sno_eff"], 0.0)
    qevap = np.where(snow, h["qflx_ev_soil"], h["qflx_liqevap_from_top_layer"])
    net = top - h["qflx_sat_excess_surf"]
    assert np.allclose(h["qflx_in_soil"], (1.0 - fh) * net - (1.0 - fsno - fh) * qevap, rtol=1e-13, atol=1e-22)
    assert np.allclose(h["qflx_top_soil_to_h2osfc"], fh * net - fh * h["qflx_ev_h2osfc"], rtol=1e-13, atol=1e-22)


@pytest.mark.parametrize("dtime", [DT, 1.0e5])
def test_overflow_and_drainage_closed_forms(dtime):
    st = _state()
    hydro = O.hydro_params(pc=0.4, mu=0.14)
    h0 = _run(st, hydro, dtime=dtime)
    h, c = _hy(st), st.col
    th = h["h2osfc_thresh"]
    fn = h["frac_h2osfc_nosnow"]
    clust = np.where(fn > 0.4, np.maximum(fn - 0.4, 0.0) ** 0.14, 0.0)
    kwet = 1.0e-4 * np.sin(np.deg2rad(h["topo_slope"]))
    over = h0 > th
    q = np.where(over, np.minimum(kwet * clust * (h0 - th), (h0 - th) / dtime), 0.0)
    q = np.where(q < np.float64(np.float32(1.0e-8)), 0.0, q)
    assert np.allclose(h["qflx_h2osfc_surf"], q, rtol=1e-13, atol=0)
    assert (h["qflx_h2osfc_surf"] > 0).any()
    if dtime > 1e4:     # the per-step cap binds for some columns
        assert np.any((q > 0) & (q == (h0 - th) / dtime))
    hp = h0 + (h["qflx_in_h2osfc"] - h["qflx_h2osfc_surf"]) * dtime
    cap = np.where(hp < 0, hp / dtime, np.minimum(c["frac_h2osfc"] * h["qinmax"], hp / dtime))
    assert np.allclose(h["qflx_h2osfc_drain"], cap, rtol=1e-12, atol=1e-18)
    assert np.all(st.col["h2osfc"] >= 0.0)


def test_special_cases():
    # no surface water fraction: rain goes to the soil side, and only the
    # infiltration excess ponds (it cannot drain through a zero fraction)
    st = _state(fh=False, deep=False)
    st.hy["frac_h2osfc_nosnow"][:] = 0.0
    _run(st, O.hydro_params())
    h = _hy(st)
    assert np.all(h["qflx_top_soil_to_h2osfc"] == 0.0)
    assert np.all(h["qflx_h2osfc_surf"] == 0.0)
    assert np.all(h["qflx_h2osfc_drain"] == 0.0)
    assert (h["qflx_infl_excess"] > 0).any()
    assert np.allclose(st.col["h2osfc"], h["qflx_infl_excess"] * DT, rtol=1e-15, atol=0)
    assert np.array_equal(h["qflx_infl"], h["qflx_in_soil"] - h["qflx_infl_excess"])
    # a pool fraction at or below pc does not overflow, however deep
    st = _state()
    st.hy["frac_h2osfc_nosnow"][:] = np.minimum(st.hy["frac_h2osfc_nosnow"], 0.4)
    _run(st, O.hydro_params(pc=0.4))
    assert np.all(st.hy["qflx_h2osfc_surf"] == 0.0)
    # nothing falling, nothing evaporating, no standing water: every flux is zero
    st = _state(deep=False)
    st.col["h2osfc"][:] = 0.0
    for k in ("qflx_rain_plus_snomelt", "qflx_liqevap_from_top_layer", "qflx_ev_soil", "qflx_ev_h2osfc"):
        st.hy[k][:] = 0.0
    _run(st, O.hydro_params())
    for k in ("qflx_infl", "qflx_surf", "qflx_in_soil", "qflx_h2osfc_drain", "qflx_infl_excess"):
        assert np.all(_hy(st)[k] == 0.0), k
    assert np.all(st.col["h2osfc"] == 0.0)


def test_filter_touches_only_listed_columns():
    st = _state(ncol=256)
    ref = st.copy()
    filt = np.arange(1, 257, 3, dtype=np.int32)
    assert O.surface_water(st, O.hydro_params(), filter_c=filt) == 0
    assert O.surface_water(ref, O.hydro_params()) == 0
    k = filt - 1
    rest = np.setdiff1d(np.arange(256), k)
    assert np.array_equal(st.col["qflx_infl"][k], ref.col["qflx_infl"][k])
    assert np.array_equal(st.col["h2osfc"][k], ref.col["h2osfc"][k])
    assert np.all(st.hy["fsat"][rest] == 0.0)


def test_patch2col_of_evaporation_fluxes():
    st = S.canopy_workload("canopy_uneven", ncol=128)
    rng = np.random.default_rng(7)
    for a in (st.sf["qflx_liqevap_from_top_layer"], st.canopy_array("qflx_ev_soil"),
              st.canopy_array("qflx_ev_h2osfc")):
        a[:] = rng.uniform(-1e-5, 1e-4, a.shape)
    active = st.canopy_array("active").copy()
    active[::5] = 0
    st.canopy_array("active")[:] = active
    O.patch2col3(st)
    col = st.pch["column"] - 1 if "column" in st.pch else None
    if col is None:
        pi, pf = st.col["patchi"], st.col["patchf"]
        col = np.repeat(np.arange(st.dims.ncol), pf - pi + 1)
    w = st.canopy_array("wtcol") * (active != 0)
    for src, dst in ((st.sf["qflx_liqevap_from_top_layer"], "qflx_liqevap_from_top_layer"),
                     (st.canopy_array("qflx_ev_soil"), "qflx_ev_soil"),
                     (st.canopy_array("qflx_ev_h2osfc"), "qflx_ev_h2osfc")):
        ref = np.zeros(st.dims.ncol)
        np.add.at(ref, col, src * w)
        assert np.allclose(st.hy[dst], ref, rtol=1e-13, atol=1e-20), dst


# ---- water tables and RenewCondensation (HydrologyNoDrainageMod.F90:355-373) ----

SAT_LEV = float(np.float32(0.9))      # `0.9` without a kind: the single-precision value
TFRZ = 273.15


def _wt_state(ncol=256, seed=0x5EED0021, snow=False):
    frac = 0.4 * S.uniform(seed, 94, np.arange(ncol, dtype=np.int64))
    st = S.soil_only(ncol, seed, snow=snow, bedrock=True, frac_h2osfc=frac)
    S.with_hydrology(st, seed)
    return st


def _lev(st):
    ns = st.dims.nlevsno
    c = st.col
    return dict(
        t=lambda k: c["t_soisno"][ns + k - 1], z=lambda k: c["z"][ns + k - 1], dz=lambda k: c["dz"][ns + k - 1],
        zi=lambda j: c["zi"][ns + j], ws=lambda k: c["watsat"][k - 1],
        liq=lambda k: c["h2osoi_liq"][ns + k - 1], ice=lambda k: c["h2osoi_ice"][ns + k - 1])


def _set_saturation(st, s_of_layer):
    """Liquid water for relative saturation s_of_layer(k) (array over columns), no ice."""
    ns = st.dims.nlevsno
    c = st.col
    for k in range(1, st.dims.nlevgrnd + 1):
        c["h2osoi_ice"][ns + k - 1] = 0.0
        if k <= st.dims.nlevsoi:
            ws = c["watsat"][k - 1]
            c["h2osoi_liq"][ns + k - 1] = s_of_layer(k) * ws * c["dz"][ns + k - 1] * 1000.0


def test_water_table_saturated_dry_and_interpolated():
    st = _wt_state()
    n = st.dims.ncol
    L = _lev(st)
    nb = st.col["nbedrock"].astype(int)
    cols = np.arange(n)
    kind = cols % 3                      # 0 saturated, 1 dry, 2 dry at layer K over saturated layers
    K = np.maximum(1, nb // 2)
    rng = np.random.default_rng(11)
    s_dry = rng.uniform(0.3, 0.8, n)
    s_wet = rng.uniform(0.95, 1.0, n)
    _set_saturation(st, lambda k: np.where(kind == 0, s_wet, np.where(kind == 1, s_dry,
                                                                       np.where(k <= K, s_dry, s_wet))))
    st.col["t_soisno"][...] = 285.0      # unfrozen: no perched table
    assert O.water_table(st) == 0
    zwt = st.hy["zwt"]
    zi = st.col["zi"]
    ns = st.dims.nlevsno
    sat = kind == 0
    assert np.array_equal(zwt[sat], zi[ns + 1, sat])                       # zi(1)
    dry = kind == 1
    assert np.array_equal(zwt[dry], zi[ns + nb[dry], cols[dry]])           # zi(nbedrock)
    mid = (kind == 2) & (K < nb) & (K > 1)
    assert mid.sum() > 20
    k = K[mid]
    c = cols[mid]
    s1 = (st.col["h2osoi_liq"][ns + k - 1, c] / (st.col["dz"][ns + k - 1, c] * 1000.0)) / st.col["watsat"][k - 1, c]
    s2 = (st.col["h2osoi_liq"][ns + k, c] / (st.col["dz"][ns + k, c] * 1000.0)) / st.col["watsat"][k, c]
    z0, z1 = st.col["z"][ns + k - 1, c], st.col["z"][ns + k, c]
    expect = np.maximum(0.0, z0 + (SAT_LEV - s1) * (z1 - z0) / (s2 - s1))
    assert np.allclose(zwt[mid], expect, rtol=1e-12, atol=1e-14)
    assert np.all((zwt[mid] >= z0 - 1e-12) & (zwt[mid] <= z1 + 1e-12))
    # volumetric water of the visited layers
    vol = st.hy["h2osoi_vol"]
    assert np.allclose(vol[nb[dry] - 1, cols[dry]],
                       st.col["h2osoi_liq"][ns + nb[dry] - 1, cols[dry]] / (st.col["dz"][ns + nb[dry] - 1, cols[dry]] * 1000.0),
                       rtol=1e-15)
    # unfrozen columns: no perched table, the frost table at zi(nlevsoi - 1)
    nl = st.dims.nlevsoi
    assert np.array_equal(st.hy["frost_table"], zi[ns + nl - 1])
    assert np.array_equal(st.hy["zwt_perched"], st.hy["frost_table"])


def test_frost_table_and_perched_water():
    st = _wt_state(seed=0x5EED0022)
    n = st.dims.ncol
    ns = st.dims.nlevsno
    nl = st.dims.nlevsoi
    cols = np.arange(n)
    rng = np.random.default_rng(12)
    kf = rng.integers(3, nl, n)                         # first frozen layer under thawed ones
    kp = np.maximum(1, kf - rng.integers(1, 3, n))      # dry layer above the perched water
    for k in range(1, st.dims.nlevgrnd + 1):
        st.col["t_soisno"][ns + k - 1] = np.where(k < kf, 280.0, 265.0)
    s_dry = rng.uniform(0.3, 0.8, n)
    s_wet = rng.uniform(0.95, 1.0, n)
    _set_saturation(st, lambda k: np.where(k <= kp, s_dry, s_wet))
    st.hy["zwt"][:] = 50.0                              # deep water table: the perched search runs
    liq = st.col["h2osoi_liq"].copy()                   # the searches see the water before condensation
    assert O.water_table(st) == 0
    zi, z = st.col["zi"], st.col["z"]
    assert np.array_equal(st.hy["frost_table"], zi[ns + kf - 1, cols])
    per = kf > kp
    k, c = kp[per], cols[per]
    assert (k == 1).any() and (k > 1).any()
    s1 = (liq[ns + k - 1, c] / (st.col["dz"][ns + k - 1, c] * 1000.0)) / st.col["watsat"][k - 1, c]
    s2 = (liq[ns + k, c] / (st.col["dz"][ns + k, c] * 1000.0)) / st.col["watsat"][k, c]
    z0, z1 = z[ns + k - 1, c], z[ns + k, c]
    expect = np.maximum(0.0, z0 + (SAT_LEV - s1) * (z1 - z0) / (s2 - s1))
    assert np.allclose(st.hy["zwt_perched"][per], expect, rtol=1e-12, atol=1e-14)
    # a water table above the frost table leaves the perched table at the frost table
    st2 = _wt_state(seed=0x5EED0022)
    st2.col["t_soisno"][...] = st.col["t_soisno"]
    st2.col["h2osoi_liq"][...] = liq
    st2.col["h2osoi_ice"][...] = 0.0
    st2.hy["zwt"][:] = 0.0
    assert O.water_table(st2) == 0
    assert np.array_equal(st2.hy["zwt_perched"], st2.hy["frost_table"])
    # frozen from the top: the frost table is the surface, nothing perched
    st3 = _wt_state(seed=0x5EED0023)
    st3.col["t_soisno"][...] = 260.0
    assert O.water_table(st3) == 0
    assert np.all(st3.hy["frost_table"] == zi[ns]) and np.all(st3.hy["zwt_perched"] == zi[ns])


def test_renew_condensation_mass_and_truncation():
    st = _wt_state(seed=0x5EED0024, snow=True)
    ns = st.dims.nlevsno
    cols = np.arange(st.dims.ncol)
    liq0 = st.col["h2osoi_liq"][ns].copy()
    ice0 = st.col["h2osoi_ice"][ns].copy()
    h = st.hy
    # one column whose sublimation removes all its ice up to a rounding residue
    bare = np.nonzero((st.col["snl"] == 0) & (st.col["frac_h2osfc"] > 0.1))[0]
    j = bare[0]
    wf_j = 1.0 - st.col["frac_h2osfc"][j]
    for ice0_j in np.linspace(0.1, 3.0, 400):
        q = ice0_j / (wf_j * DT)
        left = ice0_j - wf_j * q * DT
        if left != 0.0 and abs(left) < 1e-12 * ice0_j:
            break
    else:
        raise AssertionError("no rounding residue found")
    st.col["h2osoi_ice"][ns, j] = ice0_j
    ice0[j] = ice0_j
    h["qflx_soliddew_to_top_layer"][j] = 0.0
    h["qflx_solidevap_from_top_layer"][j] = q
    assert O.water_table(st) == 0
    snow = st.col["snl"] < 0
    assert np.array_equal(st.col["h2osoi_liq"][ns, snow], liq0[snow])     # snow columns untouched
    assert np.array_equal(st.col["h2osoi_ice"][ns, snow], ice0[snow])
    ok = ~snow
    wf = 1.0 - st.col["frac_h2osfc"]
    assert np.allclose(st.col["h2osoi_liq"][ns, ok] - liq0[ok], wf[ok] * h["qflx_liqdew_to_top_layer"][ok] * DT,
                       rtol=1e-9, atol=1e-12)
    dice = wf * (h["qflx_soliddew_to_top_layer"] - h["qflx_solidevap_from_top_layer"]) * DT
    assert np.allclose(st.col["h2osoi_ice"][ns, ok] - ice0[ok], dice[ok], rtol=1e-9, atol=1e-12)
    assert st.col["h2osoi_ice"][ns, j] == 0.0                                # the residue truncated
    assert np.all(st.col["h2osoi_ice"][ns, ok] >= 0.0)
    # sublimation beyond the ice: the reference's endrun, reported with the column
    st2 = _wt_state(seed=0x5EED0024, snow=True)
    k = bare[1]
    st2.hy["qflx_solidevap_from_top_layer"][k] = 1.0
    assert O.water_table(st2) == k + 1


# ---- drainage (HydrologyDrainageMod.F90:137-245, no aquifer layer) ----------------

WATMIN = 0.01


def _dr_state(ncol=512, seed=0x5EED0041, wet=False):
    """Soil columns whose water tables, perched tables and frost tables come
    from the water-table oracle, with ice in some layers and a few columns
    drained below the minimum water (as a Richards solve can leave them)."""
    st = _wt_state(ncol, seed=seed, snow=True)
    ns, nl = st.dims.nlevsno, st.dims.nlevsoi
    rng = np.random.default_rng(seed & 0xFFFF)
    cols = np.arange(ncol)
    kf = rng.integers(2, nl + 3, ncol)
    for k in range(1, st.dims.nlevgrnd + 1):
        st.col["t_soisno"][ns + k - 1] = np.where(k < kf, 280.0, 268.0)
        if k <= nl:
            lo, hi = (0.85, 1.0) if wet else (0.1, 1.0)
            s = rng.uniform(lo, hi, ncol)
            w = s * st.col["watsat"][k - 1] * st.col["dz"][ns + k - 1] * 1000.0
            f = np.where(k >= kf, 0.5 * rng.random(ncol), 0.0)
            st.col["h2osoi_liq"][ns + k - 1] = (1 - f) * w
            st.col["h2osoi_ice"][ns + k - 1] = f * w * 0.917
    dry = rng.random(ncol) < 0.1            # a negative layer, as left by the Richards solve
    kneg = rng.integers(1, nl + 1, ncol)
    st.col["h2osoi_liq"][ns + kneg[dry] - 1, cols[dry]] = -rng.uniform(0.0, 3.0, dry.sum())
    vol_ice = np.minimum(st.col["watsat"][:st.dims.nlevgrnd], st.col["h2osoi_ice"][ns:] / (st.col["dz"][ns:] * 917.0))
    st.col["eff_porosity"][...] = np.maximum(0.01, st.col["watsat"] - vol_ice)
    st.hy["zwt"][:] = 50.0
    for k in ("qflx_liqdew_to_top_layer", "qflx_soliddew_to_top_layer", "qflx_solidevap_from_top_layer"):
        st.hy[k][:] = 0.0
    assert O.water_table(st) == 0           # zwt, zwt_perched, frost_table from the state
    st.hy["qflx_surf"][:] = rng.uniform(0.0, 1e-4, ncol)
    return st


def _water(st):
    ns, nl = st.dims.nlevsno, st.dims.nlevsoi
    return (st.col["h2osoi_liq"][ns:ns + nl].sum(0) + st.col["h2osoi_ice"][ns:ns + nl].sum(0)
            + st.col["h2osfc"])


@pytest.mark.parametrize("wet", [False, True])
def test_drainage_water_balance_and_limits(wet):
    st = _dr_state(wet=wet)
    if not wet:                             # a column too dry to refill its bottom layer
        ns, nl = st.dims.nlevsno, st.dims.nlevsoi
        st.col["h2osoi_liq"][ns:ns + nl, 3] = 0.0
        st.col["h2osoi_liq"][ns + nl - 1, 3] = -0.5
        st.col["h2osoi_ice"][ns:ns + nl, 3] = 0.0
    ns, nl = st.dims.nlevsno, st.dims.nlevsoi
    dzmm = st.col["dz"][ns:ns + nl] * 1e3
    # a quiet column: every layer between the minimum water and saturation, no
    # ice, the water table below the bedrock, no perched water -- unchanged
    q = 5
    st.col["h2osoi_ice"][ns:ns + nl, q] = 0.0
    st.col["eff_porosity"][:nl, q] = st.col["watsat"][:nl, q]
    st.col["h2osoi_liq"][ns:ns + nl, q] = 0.5 * st.col["watsat"][:nl, q] * dzmm[:, q]
    st.hy["zwt"][q] = 100.0                 # also beyond the 80 m cap
    st.hy["frost_table"][q] = st.hy["zwt_perched"][q] = 0.5
    liq_q = st.col["h2osoi_liq"][ns:ns + nl, q].copy()
    # a top layer holding more ice than its pore space: the excess runs off
    e = 7
    st.col["h2osoi_ice"][ns, e] = 2.0 * st.col["watsat"][0, e] * dzmm[0, e]
    w0 = _water(st)
    zwt0 = st.hy["zwt"].copy()
    assert O.drainage(st, O.hydro_params()) == 0
    h = st.hy
    assert np.array_equal(st.col["h2osoi_liq"][ns:ns + nl, q], liq_q)
    assert h["zwt"][q] == 80.0 and h["qflx_drain"][q] == 0.0 and h["qflx_drain_perched"][q] == 0.0
    assert h["qflx_ice_runoff_xs"][e] > 0.0
    w1 = _water(st)
    out = (h["qflx_drain"] + h["qflx_drain_perched"] + h["qflx_ice_runoff_xs"]) * DT
    assert np.max(np.abs((w1 - w0) + out)) < 1e-10 * np.max(np.abs(w0))
    ns, nl = st.dims.nlevsno, st.dims.nlevsoi
    assert np.all(st.col["h2osoi_liq"][ns:ns + nl] >= WATMIN - 1e-12)
    assert np.all((h["zwt"] >= 0.0) & (h["zwt"] <= 80.0))
    assert np.array_equal(h["qflx_runoff"], h["qflx_drain"] + h["qflx_surf"] + h["qflx_qrgwl"]
                          + h["qflx_drain_perched"])
    assert (h["qflx_latflow_out"] > 0).any() and np.all(h["qflx_latflow_in"] == 0.0)
    if not wet:
        assert (h["qflx_rsub_sat"] < 0).any()          # the bottom layer's deficit came from outside
        assert np.any(zwt0 != h["zwt"])


def test_drainage_closed_forms():
    st = _dr_state(seed=0x5EED0042, wet=True)
    p = O.hydro_params()
    ns, nl = st.dims.nlevsno, st.dims.nlevsoi
    cols = np.arange(st.dims.ncol)
    nb = st.col["nbedrock"].astype(int)
    zi = st.col["zi"]
    zwt0, zwtp, frost = st.hy["zwt"].copy(), st.hy["zwt_perched"].copy(), st.hy["frost_table"].copy()
    ice = st.col["h2osoi_ice"][ns:ns + nl]
    dz = st.col["dz"][ns:ns + nl]
    ws = st.col["watsat"][:nl]
    assert O.drainage(st, p) == 0
    # icefrac rewritten from the ice
    icefrac = np.minimum(1.0, np.minimum(ws, ice / (dz * 917.0)) / ws)
    assert np.allclose(st.col["icefrac"][:nl], icefrac, rtol=1e-15, atol=0)
    # baseflow from the water table, with the column's ice impedance below it
    zinb = zi[ns + nb, cols]
    jwt = np.array([next((j - 1 for j in range(1, nl + 1) if zwt0[c] <= zi[ns + j, c]), nl) for c in cols])
    imped = np.empty(len(cols))
    for c in cols:
        j0 = max(jwt[c], 1)
        dzmm = dz[j0 - 1:, c] * 1e3
        imped[c] = 10.0 ** (-6.0 * np.sum(icefrac[j0 - 1:, c] * dzmm) / np.sum(dzmm))
    slope = st.hy["topo_slope"]
    q = np.where(zwt0 <= zinb, imped * p.baseflow_scalar * np.tan(np.deg2rad(slope))
                 * np.maximum(zinb - zwt0, 0.0) ** p.n_baseflow, 0.0)
    assert np.allclose(st.hy["qflx_latflow_out"], q, rtol=1e-12, atol=0)
    assert (q > 0).any() and (q == 0).any()
    # perched drainage: the transmissivity-weighted rate, never more than the layers hold
    perched = frost > zwtp
    assert perched.any()
    hk = st.col["hksat"][:nl]
    k_fr = np.array([next((k for k in range(1, nb[c] + 1) if zi[ns + k - 1, c] <= frost[c] < zi[ns + k, c]), nb[c])
                     for c in cols])
    k_pe = np.array([next((k for k in range(1, nb[c] + 1) if zi[ns + k - 1, c] <= zwtp[c] < zi[ns + k, c]), nb[c])
                     for c in cols])
    expect = np.zeros(len(cols))
    for c in np.nonzero(perched)[0]:
        ks = slice(k_pe[c] - 1, k_fr[c] - 1)
        w = dz[ks, c].sum()
        qp = (hk[ks, c] * dz[ks, c]).sum() / w if w > 0 else 0.0
        expect[c] = p.perched_baseflow_scalar * np.sin(np.deg2rad(slope[c])) * qp * (frost[c] - zwtp[c])
    got = st.hy["qflx_drain_perched"]
    assert np.all(got <= expect * (1 + 1e-12) + 1e-30)
    unlimited = np.isclose(got[perched], expect[perched], rtol=1e-12, atol=0)
    assert unlimited.mean() > 0.5, unlimited.mean()     # most columns are not limited by the layers' water
