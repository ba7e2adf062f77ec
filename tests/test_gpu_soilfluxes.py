"""SoilFluxes and the soil energy balance check on the GPU against the oracle
(bit for bit), and the full land step with them."""
import numpy as np
import pytest

from ctsm_amd import synthetic as S
from ctsm_amd.state import DeviceState
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sctx():
    from ctsm_amd.lib import Context
    c = Context()
    c.set_canopy_params(S.canopy_params(), S.pft_table())
    yield c
    c.close()


def _same(g, o, groups=("col", "pch", "can", "sf")):
    for grp in groups:
        a, b = getattr(g, grp), getattr(o, grp)
        for k in b:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"{grp}.{k}")


def _after_heat(ncol, name="canopy_uneven"):
    st = S.canopy_workload(name, ncol=ncol)
    cfg = O.canopy_config()
    O.save_before(st, cfg)
    for k, grp in (("cgrnd", st.pch), ("cgrnds", st.can), ("cgrndl", st.can)):
        grp[k][:] = 0.0
    rc, bad = O.canopyfluxes(st, cfg, S.canopy_params(), O.PftStruct(S.pft_table()))
    assert rc == 0, bad
    rc, bad = O.soiltemperature(st, cfg)
    assert rc == 0, bad
    return st, cfg


def test_soilfluxes_dropin_bitwise(sctx):
    st, cfg = _after_heat(384)
    print(f"case: SoilFluxes drop-in, 384 columns / {st.dims.npatch} patches", flush=True)
    g, o = st.copy(), st.copy()
    sctx.soilfluxes(g)
    assert O.soilfluxes(o, cfg) == 0
    _same(g, o)
    assert np.max(np.abs(g.sf["errsoi_col"])) < 1e-7


def test_soilfluxes_dropin_filters(sctx):
    st, cfg = _after_heat(256)
    fc = np.arange(st.dims.ncol, 0, -3, dtype=np.int32)
    cols = set(fc.tolist())
    fp = np.array([p + 1 for p in range(st.dims.npatch) if st.pch["column"][p] in cols], dtype=np.int32)[::-1].copy()
    print(f"case: SoilFluxes drop-in, {len(fc)} columns / {len(fp)} patches by filter", flush=True)
    g, o = st.copy(), st.copy()
    sctx.soilfluxes(g, filter_c=fc, filter_p=fp)
    assert O.soilfluxes(o, cfg, filter_c=fc, filter_p=fp) == 0
    _same(g, o)


def test_soilfluxes_dev_and_balance_check(sctx):
    st, cfg = _after_heat(512)
    print(f"case: SoilFluxes dev + balance check, 512 columns / {st.dims.npatch} patches", flush=True)
    ds = DeviceState(st, sctx)
    try:
        sctx.soilfluxes_dev(ds)
        m, idx = sctx.soil_balance_check_dev(ds)
        sctx.sync()
        g = ds.to_host()
    finally:
        ds.free()
    o = st.copy()
    assert O.soilfluxes(o, cfg) == 0
    _same(g, o)
    assert (m, idx) == O.soil_balance(o)
    # planted maximum with a tie: the lowest column wins
    o.sf["errsoi_col"][[40, 7]] = 3.0
    assert sctx.soil_balance_check(o) == (3.0, 8)
    assert O.soil_balance(o) == (3.0, 8)


@pytest.mark.parametrize("nsteps,name", [(1, "canopy"), (3, "canopy_uneven")])
def test_land_step_full_bitwise(sctx, nsteps, name):
    st = S.canopy_workload(name, ncol=384)
    print(f"case: full land step x{nsteps}, {name}, 384 columns / {st.dims.npatch} patches", flush=True)
    ds = DeviceState(st, sctx)
    worst = 0.0
    try:
        for _ in range(nsteps):
            sctx.land_step_full_dev(ds)
            m, _ = sctx.soil_balance_check_dev(ds)
            worst = max(worst, m)
        sctx.sync()
        g = ds.to_host()
    finally:
        ds.free()
    o = st.copy()
    cfg = O.canopy_config()
    pft = O.PftStruct(S.pft_table())
    for _ in range(nsteps):
        rc, bad = O.land_step_full(o, cfg, S.canopy_params(), pft)
        assert rc == 0, bad
    _same(g, o)
    assert worst < 1e-7, worst             # the reference warns at 1e-5 W/m2


def test_energy_balance_check_dev(sctx):
    """EnergyBalanceCheck's patch errors on the device after full land steps:
    per-patch errseb / errlon bit for bit with the oracle's on the same
    state, through a descending patch filter with inactive patches, the max
    and its patch as the oracle finds them, and both closed below 1e-9 W/m2."""
    st = S.canopy_workload("canopy_uneven", ncol=700)
    st.can["active"][::9] = 0
    print(f"case: energy balance check after 2 full land steps, 700 columns / {st.dims.npatch} patches", flush=True)
    hydro = S.hydro_params()
    ds = DeviceState(st, sctx)
    filt = np.arange(st.dims.npatch, 0, -1, dtype=np.int32)
    fdev = sctx.alloc(filt.nbytes)
    seb_dev = sctx.alloc(8 * len(filt))
    lon_dev = sctx.alloc(8 * len(filt))
    try:
        for _ in range(2):
            sctx.land_step_full_dev(ds, hydro=hydro, water_table=True, drainage=True)
        sctx.h2d(fdev, filt)
        (ms, ps), (ml, pl) = sctx.energy_balance_check_dev(ds, filterp_dev=fdev, n=len(filt), errseb_dev=seb_dev,
                                                           errlon_dev=lon_dev)
        seb, lon = np.empty(len(filt)), np.empty(len(filt))
        sctx.d2h(seb, seb_dev)
        sctx.d2h(lon, lon_dev)
        g = ds.to_host()
    finally:
        for p in (fdev, seb_dev, lon_dev):
            sctx.free(p)
        ds.free()
    o_seb, o_lon, (oms, ops), (oml, opl) = O.energy_balance(g, filter_p=filt)
    np.testing.assert_array_equal(seb, o_seb)
    np.testing.assert_array_equal(lon, o_lon)
    assert (ms, ps) == (oms, ops) and (ml, pl) == (oml, opl)
    assert ms < 1e-9 and ml < 1e-9, (ms, ml)
