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
