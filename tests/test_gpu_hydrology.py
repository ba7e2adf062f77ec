"""The surface water input chain on the GPU against the oracle, bit for bit:
the drop-in entry on host arrays (both h2osfcflag values, with a filter),
the device-resident entry, and the full land step with the chain feeding
SoilWater's infiltration."""
import numpy as np
import pytest

from ctsm_amd import synthetic as S
from ctsm_amd.state import DeviceState
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DT = 1800.0


@pytest.fixture(scope="module")
def hctx():
    from ctsm_amd.lib import Context
    c = Context()
    c.set_canopy_params(S.canopy_params(), S.pft_table())
    yield c
    c.close()


def _state(ncol, seed=0x5EED0011):
    cols = np.arange(ncol, dtype=np.int64)
    st = S.soil_only(ncol, seed, snow=True, bedrock=True, frac_h2osfc=0.6 * S.uniform(seed, 90, cols))
    S.with_hydrology(st, seed)
    st.col["h2osfc"][:] = 100.0 * S.uniform(seed, 91, cols)
    for j in range(3):
        st.col["icefrac"][j] = np.where(S.uniform(seed, 92, cols, j) < 0.3, S.uniform(seed, 93, cols, j), 0.0)
    return st


def _same(g, o, groups=("col", "hy")):
    for grp in groups:
        a, b = getattr(g, grp), getattr(o, grp)
        for k in b:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{grp}.{k}")


@pytest.mark.parametrize("flag,dtime", [(1, DT), (0, DT), (1, 1.0e5)])
def test_surface_water_dropin_bitwise(hctx, flag, dtime):
    st = _state(1000)
    print(f"case: surface water drop-in, 1000 columns, h2osfcflag={flag}, dtime={dtime}", flush=True)
    hydro = O.hydro_params(h2osfcflag=flag)
    g, o = st.copy(), st.copy()
    hctx.surface_water(g, hydro, dtime=dtime)
    assert O.surface_water(o, hydro, dtime=dtime) == 0
    _same(g, o)
    assert (o.hy["qflx_infl_excess"] > 0).any()
    assert (o.hy["qflx_h2osfc_surf"] > 0).any() == (flag == 1)     # no overflow without the pool


def test_surface_water_dropin_filter(hctx):
    st = _state(600)
    fc = np.arange(st.dims.ncol, 0, -4, dtype=np.int32)
    print(f"case: surface water drop-in, {len(fc)} of 600 columns by filter", flush=True)
    hydro = O.hydro_params()
    g, o = st.copy(), st.copy()
    hctx.surface_water(g, hydro, filter_c=fc)
    assert O.surface_water(o, hydro, filter_c=fc) == 0
    _same(g, o)


def test_surface_water_dev_bitwise(hctx):
    st = _state(4096)
    perm = (np.random.default_rng(3).permutation(st.dims.ncol) + 1).astype(np.int32)
    print("case: surface water device entry, 4096 columns, permuted filter", flush=True)
    hydro = O.hydro_params()
    ds = DeviceState(st, hctx)
    f = hctx.device_array(perm)
    try:
        hctx.surface_water_dev(ds, hydro, filter_dev=f, n=len(perm))
        hctx.sync()
        g = ds.to_host()
    finally:
        hctx.free(f)
        ds.free()
    o = st.copy()
    assert O.surface_water(o, hydro, filter_c=perm) == 0
    _same(g, o)


@pytest.mark.parametrize("nsteps,name", [(1, "canopy"), (3, "canopy_uneven")])
def test_land_step_full_with_surface_water_bitwise(hctx, nsteps, name):
    st = S.canopy_workload(name, ncol=384)
    print(f"case: full land step + surface water x{nsteps}, {name}, 384 columns / {st.dims.npatch} patches",
          flush=True)
    hydro = O.hydro_params()
    ds = DeviceState(st, hctx)
    worst = 0.0
    try:
        for _ in range(nsteps):
            hctx.land_step_full_dev(ds, hydro=hydro)
            m, _ = hctx.soil_balance_check_dev(ds)
            worst = max(worst, m)
        hctx.sync()
        g = ds.to_host()
    finally:
        ds.free()
    o = st.copy()
    cfg = O.canopy_config()
    pft = O.PftStruct(S.pft_table())
    for _ in range(nsteps):
        rc, bad = O.land_step_full(o, cfg, S.canopy_params(), pft, hydro=hydro)
        assert rc == 0, bad
    _same(g, o, groups=("col", "pch", "can", "sf", "hy"))
    assert worst < 1e-7, worst
    # the chain fed SoilWater: infiltration is no longer the synthetic input
    assert np.any(o.col["qflx_infl"] != st.col["qflx_infl"])


# ---- water tables and RenewCondensation ----------------------------------------

def _wt_state(ncol, seed=0x5EED0031):
    """Columns with random freezing profiles, saturations and water tables, so
    every branch of both searches is taken."""
    cols = np.arange(ncol, dtype=np.int64)
    st = S.soil_only(ncol, seed, snow=True, bedrock=True, frac_h2osfc=0.4 * S.uniform(seed, 94, cols))
    S.with_hydrology(st, seed)
    ns, nl = st.dims.nlevsno, st.dims.nlevsoi
    rng = np.random.default_rng(seed & 0xFFFF)
    kf = rng.integers(1, nl + 2, ncol)                   # nl + 1: unfrozen column
    for k in range(1, st.dims.nlevgrnd + 1):
        st.col["t_soisno"][ns + k - 1] = np.where(k < kf, 275.0 + 10 * rng.random(ncol), 273.15 - 8 * rng.random(ncol))
        if k <= nl:
            s = np.where(rng.random(ncol) < 0.5, rng.uniform(0.2, 0.9, ncol), rng.uniform(0.9, 1.0, ncol))
            w = s * st.col["watsat"][k - 1] * st.col["dz"][ns + k - 1] * 1000.0
            f = np.where(k >= kf, rng.random(ncol), 0.0)
            st.col["h2osoi_liq"][ns + k - 1] = (1 - f) * w
            st.col["h2osoi_ice"][ns + k - 1] = f * w * 0.917
    st.hy["zwt"][:] = rng.uniform(0.0, 10.0, ncol)
    ice1 = st.col["h2osoi_ice"][ns]
    st.hy["qflx_solidevap_from_top_layer"][:] = rng.random(ncol) * (st.hy["qflx_soliddew_to_top_layer"] + ice1 / DT)
    return st


def test_water_table_dropin_bitwise(hctx):
    st = _wt_state(2000)
    print("case: water tables + RenewCondensation drop-in, 2000 columns", flush=True)
    g, o = st.copy(), st.copy()
    hctx.water_table(g)
    assert O.water_table(o) == 0
    _same(g, o)
    # every branch ran: perched tables below the frost table, interpolated and bedrock tables
    assert np.any(o.hy["zwt_perched"] != o.hy["frost_table"])
    assert len(np.unique(o.hy["zwt"])) > 500


def test_water_table_dev_bitwise(hctx):
    st = _wt_state(4096, seed=0x5EED0032)
    perm = (np.random.default_rng(5).permutation(st.dims.ncol) + 1).astype(np.int32)[:3000]
    print("case: water tables device entry, 3000 of 4096 columns, permuted filter", flush=True)
    ds = DeviceState(st, hctx)
    f = hctx.device_array(perm)
    try:
        hctx.water_table_dev(ds, filter_dev=f, n=len(perm))
        hctx.sync()
        g = ds.to_host()
    finally:
        hctx.free(f)
        ds.free()
    o = st.copy()
    assert O.water_table(o, filter_c=perm) == 0
    _same(g, o)


@pytest.mark.parametrize("nsteps,name", [(1, "canopy"), (3, "canopy_uneven")])
def test_land_step_full_with_water_tables_bitwise(hctx, nsteps, name):
    st = S.canopy_workload(name, ncol=320, seed=0xC75D0031)
    print(f"case: full land step + surface water + water tables x{nsteps}, {name}, 320 columns / "
          f"{st.dims.npatch} patches", flush=True)
    hydro = O.hydro_params()
    ds = DeviceState(st, hctx)
    try:
        for _ in range(nsteps):
            hctx.land_step_full_dev(ds, hydro=hydro, water_table=True)
        hctx.sync()
        g = ds.to_host()
    finally:
        ds.free()
    o = st.copy()
    cfg = O.canopy_config()
    pft = O.PftStruct(S.pft_table())
    for _ in range(nsteps):
        rc, bad = O.land_step_full(o, cfg, S.canopy_params(), pft, hydro=hydro, water_table_after=True)
        assert rc == 0, bad
    _same(g, o, groups=("col", "pch", "can", "sf", "hy"))
    # the loop is closed: the water tables the next step's fsat reads were recomputed
    assert np.any(o.hy["zwt"] != st.hy["zwt"]) and np.any(o.hy["frost_table"] != st.hy["frost_table"])


def test_water_table_endrun_reports_column():
    from ctsm_amd.lib import Context, CtsmError
    st = _wt_state(64, seed=0x5EED0033)
    bare = np.nonzero(st.col["snl"] == 0)[0]
    k = int(bare[3])
    st.hy["qflx_solidevap_from_top_layer"][k] = 1.0     # far more than the top layer's ice
    print(f"case: RenewCondensation sublimation beyond the ice, column {k + 1} of 64", flush=True)
    assert O.water_table(st.copy()) == k + 1
    c = Context()
    try:
        with pytest.raises(CtsmError) as ei:
            c.water_table(st)
        assert ei.value.index == k + 1
    finally:
        c.close()


# ---- drainage --------------------------------------------------------------------

def _dr_state(ncol, seed=0x5EED0061):
    from tests.test_oracle_hydrology import _dr_state as make
    return make(ncol, seed=seed)


def test_drainage_dropin_bitwise(hctx):
    st = _dr_state(1500)
    print("case: drainage drop-in, 1500 columns", flush=True)
    hydro = O.hydro_params()
    g, o = st.copy(), st.copy()
    hctx.drainage(g, hydro)
    assert O.drainage(o, hydro) == 0
    _same(g, o)
    assert (o.hy["qflx_drain_perched"] > 0).any() and (o.hy["qflx_latflow_out"] > 0).any()


def test_drainage_dev_bitwise(hctx):
    st = _dr_state(4096, seed=0x5EED0062)
    perm = (np.random.default_rng(9).permutation(st.dims.ncol) + 1).astype(np.int32)[:3500]
    print("case: drainage device entry, 3500 of 4096 columns, permuted filter", flush=True)
    hydro = O.hydro_params()
    ds = DeviceState(st, hctx)
    f = hctx.device_array(perm)
    try:
        hctx.drainage_dev(ds, hydro, filter_dev=f, n=len(perm))
        hctx.sync()
        g = ds.to_host()
    finally:
        hctx.free(f)
        ds.free()
    o = st.copy()
    assert O.drainage(o, hydro, filter_c=perm) == 0
    _same(g, o)


@pytest.mark.parametrize("nsteps,name", [(1, "canopy"), (4, "canopy_uneven")])
def test_land_step_full_with_drainage_bitwise(hctx, nsteps, name):
    st = S.canopy_workload(name, ncol=320, seed=0xC75D0061)
    print(f"case: full land step + surface water + water tables + drainage x{nsteps}, {name}, 320 columns / "
          f"{st.dims.npatch} patches", flush=True)
    hydro = O.hydro_params()
    ds = DeviceState(st, hctx)
    worst = 0.0
    try:
        for _ in range(nsteps):
            hctx.land_step_full_dev(ds, hydro=hydro, water_table=True, drainage=True)
            m, _ = hctx.soil_balance_check_dev(ds)
            worst = max(worst, m)
        hctx.sync()
        g = ds.to_host()
    finally:
        ds.free()
    o = st.copy()
    cfg = O.canopy_config()
    pft = O.PftStruct(S.pft_table())
    for _ in range(nsteps):
        rc, bad = O.land_step_full(o, cfg, S.canopy_params(), pft, hydro=hydro, water_table_after=True,
                                   drainage_after=True)
        assert rc == 0, bad
    _same(g, o, groups=("col", "pch", "can", "sf", "hy"))
    assert worst < 1e-7, worst
    ns, nl = st.dims.nlevsno, st.dims.nlevsoi
    assert np.all(o.col["h2osoi_liq"][ns:ns + nl] >= 0.01 - 1e-12)     # the minimum-water corrections ran


def test_full_land_step_48_steps_bitwise(hctx):
    """One day (48 steps) of the full land step with every hydrology piece, on
    snow and bedrock mixes: the GPU stays bit for bit on the oracle, the energy
    balance closes every step and no soil layer falls below the minimum water."""
    st = S.canopy_workload("canopy_uneven", ncol=96, seed=0xC75D0091)
    print(f"case: 48 full land steps with hydrology and drainage, 96 columns / {st.dims.npatch} patches",
          flush=True)
    hydro = O.hydro_params()
    ds = DeviceState(st, hctx)
    worst = 0.0
    try:
        for _ in range(48):
            hctx.land_step_full_dev(ds, hydro=hydro, water_table=True, drainage=True)
            m, _ = hctx.soil_balance_check_dev(ds)
            worst = max(worst, m)
        hctx.sync()
        g = ds.to_host()
    finally:
        ds.free()
    o = st.copy()
    cfg = O.canopy_config()
    pft = O.PftStruct(S.pft_table())
    for _ in range(48):
        rc, bad = O.land_step_full(o, cfg, S.canopy_params(), pft, hydro=hydro, water_table_after=True,
                                   drainage_after=True)
        assert rc == 0, bad
    _same(g, o, groups=("col", "pch", "can", "sf", "hy"))
    assert worst < 1e-7, worst
    ns, nl = st.dims.nlevsno, st.dims.nlevsoi
    assert np.all(o.col["h2osoi_liq"][ns:ns + nl] >= 0.01 - 1e-12)


def test_hydrology_entries_reject_bad_arguments(hctx):
    """Filters outside the allocation and incomplete states are refused with
    CTSM_E_ARG before anything runs; the context stays usable."""
    from ctsm_amd.lib import CTSM_E_ARG, CtsmError
    st = _dr_state(64, seed=0x5EED0063)
    print("case: hydrology entries with bad arguments, 64 columns", flush=True)
    hydro = O.hydro_params()
    bad = np.array([1, 65], dtype=np.int32)
    for call in (lambda: hctx.surface_water(st, hydro, filter_c=bad),
                 lambda: hctx.water_table(st, filter_c=bad),
                 lambda: hctx.drainage(st, hydro, filter_c=bad)):
        with pytest.raises(CtsmError) as ei:
            call()
        assert ei.value.code == CTSM_E_ARG
    # a drainage state without its outputs
    import ctypes as C
    cs, _ = st.structs()
    dr = st.drainage_struct()
    dr.qflx_runoff = None
    rc = hctx.L.ctsm_drainage(hctx.h, C.byref(st.layout()), C.byref(cs), C.byref(dr), C.byref(hydro),
                              st.dims.ncol, None, 1800.0)
    assert rc == CTSM_E_ARG
    g, o = st.copy(), st.copy()
    hctx.drainage(g, hydro)
    assert O.drainage(o, hydro) == 0
    _same(g, o)
