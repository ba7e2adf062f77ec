"""The patch-to-column sums that stage a block's patches in LDS (the
transpiration sink and the column evaporation / dew fluxes, taken when no
filter is given) against the oracle, bit for bit: blocks that fit the staging
space and blocks that fall back to per-lane reads (more than 16 patches per
column on average), with inactive patches and a partial last block."""
import numpy as np
import pytest

from ctsm_amd import synthetic as S
from ctsm_amd.state import DeviceState
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pmin,pmax", [(1, 16), (20, 40)])
def test_staged_patch_sums_bitwise(pmin, pmax):
    from ctsm_amd.lib import Context
    c = Context()
    c.set_canopy_params(S.canopy_params(), S.pft_table())
    seed = 0xC75D0071
    st = S.with_canopy(S.soil_only(200, seed, snow=True, bedrock=True), seed, pmax=pmax, pmin=pmin)
    act = st.canopy_array("active")
    act[::7] = 0
    print(f"case: staged patch sums, 200 columns / {st.dims.npatch} patches ({pmin}-{pmax} per column)", flush=True)
    hydro = O.hydro_params()
    ds = DeviceState(st, c)
    try:
        for _ in range(2):
            c.land_step_full_dev(ds, hydro=hydro, water_table=True, drainage=True)
        c.sync()
        g = ds.to_host()
    finally:
        ds.free()
        c.close()
    o = st.copy()
    cfg = O.canopy_config()
    pft = O.PftStruct(S.pft_table())
    for _ in range(2):
        rc, bad = O.land_step_full(o, cfg, S.canopy_params(), pft, hydro=hydro, water_table_after=True,
                                   drainage_after=True)
        assert rc == 0, bad
    for grp in ("col", "pch", "can", "sf", "hy"):
        for k, b in getattr(o, grp).items():
            np.testing.assert_array_equal(getattr(g, grp)[k], b, err_msg=f"{grp}.{k}")
    # the sums were not trivial
    assert np.any(o.can["qflx_tran_veg_col"] != 0.0) and np.any(o.hy["qflx_ev_soil"] != 0.0)
