"""The column water mass (ctsm_water_mass_dev), the column water balance check
(ctsm_water_balance_check_dev) and the solar radiation balance check
(ctsm_solar_balance_check_dev) on the GPU against the oracle, bit for bit
(TotalWaterAndHeatMod.F90:92-393, BalanceCheckMod.F90:588-630, 964-1012):
with and without excess ice, with and without the dynbal baselines, all
columns and a permuted filter, the mass before and after a full land step,
max / maxloc, and the argument checks.

Each test prints its case before the GPU work starts.
"""
import numpy as np
import pytest

from ctsm_amd import balance as B
from ctsm_amd import synthetic as S
from ctsm_amd.state import DeviceState
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bctx():
    from ctsm_amd.lib import Context, default_config
    ctx = Context(default_config())
    ctx.set_canopy_params(S.canopy_params(), S.pft_table())
    yield ctx
    ctx.close()


def _workload(ncol, seed, excess_ice=True):
    st = S.canopy_workload("canopy_uneven", ncol=ncol)
    rng = np.random.default_rng(seed)
    st.can["snocan"][:] = rng.uniform(0.0, 0.3, st.dims.npatch) * (rng.random(st.dims.npatch) < 0.5)
    st.can["active"][rng.random(st.dims.npatch) < 0.05] = 0
    return st, B.synthetic(st, seed + 1, excess_ice=excess_ice)


def _upload(ctx, arr):
    p = ctx.alloc(max(arr.nbytes, 8))
    ctx.h2d(p, np.ascontiguousarray(arr))
    return p


def _mass_gpu(ctx, ds, dwb, ncol, filt=None, subtract=False):
    out = _upload(ctx, np.zeros(ncol))
    fd = _upload(ctx, filt) if filt is not None else None
    try:
        ctx.water_mass_dev(ds, dwb, out, filter_dev=fd, n=None if filt is None else len(filt), subtract=subtract)
        got = np.empty(ncol)
        ctx.sync()
        ctx.d2h(got, out)
        return got
    finally:
        ctx.free(out)
        if fd is not None:
            ctx.free(fd)


@pytest.mark.parametrize("excess_ice,subtract,filtered", [(True, False, False), (False, True, True),
                                                          (True, True, True)],
                         ids=["excess-all", "noexcess-baselines-filter", "excess-baselines-filter"])
def test_water_mass_bitwise(bctx, excess_ice, subtract, filtered):
    ncol = 3000
    st, wb = _workload(ncol, 11 + 2 * excess_ice + subtract, excess_ice)
    filt = None
    if filtered:
        filt = np.random.default_rng(3).permutation(np.arange(1, ncol + 1))[: ncol * 2 // 3].astype(np.int32)
    print(f"case: water mass excess_ice={excess_ice} subtract={subtract} filter={filtered}, {ncol} columns / "
          f"{st.dims.npatch} patches", flush=True)
    ds = DeviceState(st, bctx)
    dwb = B.DeviceWaterBalance(bctx, wb)
    try:
        got = _mass_gpu(bctx, ds, dwb, ncol, filt, subtract)
    finally:
        dwb.free()
        ds.free()
    ref = O.water_mass(st, wb, filter_c=filt, subtract=subtract)
    if filt is not None:
        np.testing.assert_array_equal(got[filt - 1], ref[filt - 1])
        rest = np.setdiff1d(np.arange(ncol), filt - 1)
        assert np.all(got[rest] == 0.0)
    else:
        np.testing.assert_array_equal(got, ref)
    assert np.all(np.isfinite(ref)) and np.ptp(ref) > 100.0


def test_water_balance_over_a_land_step_bitwise(bctx):
    """begwb on the state, a full land step on the device, endwb, then the
    check: the masses and errh2o equal the oracle's, and max / maxloc agree."""
    ncol = 2048
    st, wb = _workload(ncol, 31)
    print(f"case: water balance over a land step, {ncol} columns / {st.dims.npatch} patches", flush=True)
    ds = DeviceState(st, bctx)
    dwb = B.DeviceWaterBalance(bctx, wb)
    try:
        begwb = _mass_gpu(bctx, ds, dwb, ncol)
        bctx.land_step_dev(ds)
        bctx.sync()
        endwb = _mass_gpu(bctx, ds, dwb, ncol)
        after = ds.to_host()
    finally:
        dwb.free()
        ds.free()
    np.testing.assert_array_equal(begwb, O.water_mass(st, wb))
    np.testing.assert_array_equal(endwb, O.water_mass(after, wb))
    assert np.any(begwb != endwb)
    w2 = wb.copy()
    w2["begwb"][:] = begwb
    w2["endwb"][:] = endwb
    dw2 = B.DeviceWaterBalance(bctx, w2)
    ds2 = DeviceState(after, bctx)
    try:
        m, c = bctx.water_balance_check_dev(ds2, dw2)
        err = dw2.fetch("errh2o")
    finally:
        dw2.free()
        ds2.free()
    om, oc = O.water_balance(after, w2)
    np.testing.assert_array_equal(err, w2["errh2o"])
    assert (m, c) == (om, oc) and m > 0.0


def test_water_balance_filter_ties_and_inactive(bctx):
    ncol = 777
    st, wb = _workload(ncol, 41)
    wb["begwb"][:] = np.round(wb["begwb"])                 # whole numbers: begwb + 4 is exact
    wb["endwb"][:] = wb["begwb"]
    for f in B.FLUXES:
        wb[f][:] = 0.0
    wb["col_active"][:] = 1
    wb["col_active"][[100, 5]] = 0
    wb["endwb"][[300, 200, 100, 5]] += 4.0                 # a tie: the lowest active column is 200
    filt = np.array([700, 301, 201, 101, 6, 2], dtype=np.int32)
    print(f"case: water balance ties / inactive / filter, {ncol} columns", flush=True)
    ds = DeviceState(st, bctx)
    dwb = B.DeviceWaterBalance(bctx, wb)
    fd = _upload(bctx, filt)
    try:
        m_all, c_all = bctx.water_balance_check_dev(ds, dwb)
        m_f, c_f = bctx.water_balance_check_dev(ds, dwb, filter_dev=fd, n=len(filt))
        m0, c0 = bctx.water_balance_check_dev(ds, dwb, filter_dev=fd, n=0)
    finally:
        bctx.free(fd)
        dwb.free()
        ds.free()
    assert (m_all, c_all) == O.water_balance(st, wb.copy()) == (4.0, 201)
    assert (m_f, c_f) == O.water_balance(st, wb.copy(), filter_c=filt) == (4.0, 201)
    assert (m0, c0) == (0.0, 0)


def test_solar_balance_bitwise(bctx):
    ncol = 1500
    st, _ = _workload(ncol, 51)
    rng = np.random.default_rng(52)
    n, npch = st.dims.ncol, st.dims.npatch
    ng = st.dims.ngrid
    assert ng > 0
    solad = rng.uniform(0.0, 400.0, (2, n))
    solai = rng.uniform(0.0, 150.0, (2, ng))
    fsa = rng.uniform(0.0, 900.0, npch)
    fsr = rng.uniform(0.0, 200.0, npch)
    filt = np.random.default_rng(53).permutation(np.arange(1, npch + 1))[: npch // 2].astype(np.int32)
    print(f"case: solar balance, {ncol} columns / {npch} patches, all and a permuted half", flush=True)
    ds = DeviceState(st, bctx)
    ptrs = [_upload(bctx, a) for a in (fsa, fsr, solad, solai)]
    out = _upload(bctx, np.zeros(npch))
    fd = _upload(bctx, filt)
    try:
        m, p = bctx.solar_balance_check_dev(ds, *ptrs, errsol_dev=out)
        got = np.empty(npch)
        bctx.d2h(got, out)
        mf, pf = bctx.solar_balance_check_dev(ds, *ptrs, filterp_dev=fd, n=len(filt), errsol_dev=out)
        gotf = np.empty(npch)
        bctx.d2h(gotf, out)
    finally:
        for q in ptrs + [out, fd]:
            bctx.free(q)
        ds.free()
    err, (om, op) = O.solar_balance(st, fsa, fsr, solad, solai)
    np.testing.assert_array_equal(got, err)
    assert (m, p) == (om, op)
    errf, (omf, opf) = O.solar_balance(st, fsa, fsr, solad, solai, filter_p=filt)
    np.testing.assert_array_equal(gotf[: len(filt)], errf)
    assert (mf, pf) == (omf, opf)


def test_balance_entries_reject_bad_arguments(bctx):
    from ctsm_amd.lib import CTSM_E_ARG, CtsmError
    ncol = 64
    st, wb = _workload(ncol, 61)
    print(f"case: water / solar balance argument checks, {ncol} columns", flush=True)
    ds = DeviceState(st, bctx)
    dwb = B.DeviceWaterBalance(bctx, wb)
    bad = _upload(bctx, np.array([1, ncol + 5], dtype=np.int32))
    out = _upload(bctx, np.zeros(ncol))
    try:
        with pytest.raises(CtsmError) as ei:                 # a column outside the layout
            bctx.water_balance_check_dev(ds, dwb, filter_dev=bad, n=2)
        assert ei.value.code == CTSM_E_ARG
        with pytest.raises(CtsmError) as ei:
            bctx.water_mass_dev(ds, dwb, out, filter_dev=bad, n=2)
            bctx.water_balance_check_dev(ds, dwb)          # the mass is enqueued; a waiting call reports it
        assert ei.value.code == CTSM_E_ARG
        w2 = B.WaterBalance({k: v for k, v in wb.a.items() if k != "qflx_drain"})
        dw2 = B.DeviceWaterBalance(bctx, w2)
        try:
            with pytest.raises(CtsmError) as ei:             # a flux missing
                bctx.water_balance_check_dev(ds, dw2)
            assert ei.value.code == CTSM_E_ARG
        finally:
            dw2.free()
        # the context still works after the reported errors
        m, c = bctx.water_balance_check_dev(ds, dwb)
        assert (m, c) == O.water_balance(st, wb.copy())
    finally:
        bctx.free(bad)
        bctx.free(out)
        dwb.free()
        ds.free()
