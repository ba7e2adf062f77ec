"""Plant hydraulic stress on the GPU against the oracle, bit for bit: the
root-soil conductances of PhotosynthesisHydraulicStress, the PHS vertical
transpiration sink and calcstress (the vegetation water potential's Newton
solve), through the drop-in entries (host arrays) and the device-resident ones,
with and without filters, chained as CTSM calls them; the vulnerability curve
against the reference's known answer; error paths."""
import numpy as np
import pytest

from ctsm_amd import phs as PH
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CONDUCTANCE = ("k_soil_root", "root_conductance", "soil_conductance")
SINK = ("qflx_rootsoi", "qflx_phs_neg", "qflx_hydr_redist")


@pytest.fixture(scope="module")
def ctx():
    from ctsm_amd.lib import Context
    c = Context()
    c.set_phs_params(PH.default_params())
    yield c
    c.close()


def _same(g, o, names):
    for k in names:
        np.testing.assert_array_equal(g[k], o[k], err_msg=k)


@pytest.mark.parametrize("ncol,seed", [(1, 1), (257, 2), (20000, 3)])
def test_conductance_dropin_all_patches(ctx, ncol, seed):
    A = PH.synthetic(ncol, seed=seed)
    ref = A.copy()
    assert O.phs_soil_root_conductance(ref, PH.default_params()) == 0
    ctx.phs_soil_root_conductance(A.layout(), A.host_struct(), A.npatch)
    _same(A, ref, CONDUCTANCE)


def test_conductance_dev_with_filter(ctx):
    A = PH.synthetic(3000, seed=7)
    rng = np.random.default_rng(8)
    filt = np.sort(rng.choice(np.arange(1, A.npatch + 1), A.npatch // 3, replace=False)).astype(np.int32)
    ref = A.copy()
    assert O.phs_soil_root_conductance(ref, PH.default_params(), filt) == 0
    D = PH.DevicePhs(ctx, A)
    fd = ctx.alloc(filt.nbytes)
    ctx.h2d(fd, filt)
    ctx.phs_soil_root_conductance(A.layout(), D.struct(), len(filt), fd, dev=True)
    out = D.fetch()
    _same(out, ref, CONDUCTANCE)
    # patches outside the filter keep their values
    rest = np.setdiff1d(np.arange(A.npatch), filt - 1)
    np.testing.assert_array_equal(out["k_soil_root"][:, rest], A["k_soil_root"][:, rest])
    ctx.free(fd)
    D.free()


@pytest.mark.parametrize("ncol,seed", [(1, 11), (999, 12), (50000, 13)])
def test_root_sink_dropin(ctx, ncol, seed):
    A = PH.synthetic(ncol, seed=seed)
    ref = A.copy()
    assert O.phs_root_sink(ref) == 0
    ctx.phs_root_sink(A.layout(), A.host_struct(), A.ncol)
    _same(A, ref, SINK)


def test_chain_on_device_with_column_filter(ctx):
    # conductances for every patch, then the sink over a column filter, as
    # CanopyFluxes and then SoilWater would call them
    A = PH.synthetic(4096, seed=21)
    rng = np.random.default_rng(22)
    fc = np.sort(rng.choice(np.arange(1, A.ncol + 1), 3000, replace=False)).astype(np.int32)
    ref = A.copy()
    assert O.phs_soil_root_conductance(ref, PH.default_params()) == 0
    assert O.phs_root_sink(ref, fc) == 0
    D = PH.DevicePhs(ctx, A)
    fd = ctx.alloc(fc.nbytes)
    ctx.h2d(fd, fc)
    ctx.phs_soil_root_conductance(A.layout(), D.struct(), A.npatch, None, dev=True)
    ctx.phs_root_sink(A.layout(), D.struct(), len(fc), fd, dev=True)
    out = D.fetch()
    _same(out, ref, CONDUCTANCE + SINK)
    ctx.free(fd)
    D.free()


def test_vulnerability_curve_known_answer(ctx):
    # rai = krmax = 0.25 + z = 1 in layer 2, so root_conductance there is plc(smp)
    # of the root segment: test_Photosynthesis.pf:67 (psi50 -150000 mm, ck 3.95)
    from ctsm_amd.lib import Context
    c = Context()
    try:
        c.set_phs_params(PH.phs_params([1.0], [0.29e-3], [0.31e6], [1.0], np.full((4, 1), -150000.0),
                                       np.full((4, 1), 3.95)))
        A = PH.PhsArrays(1, 1)
        A["column"][:] = 1
        A["patchi"][:] = 1
        A["patchf"][:] = 1
        A["dz"][12:] = 0.1
        A["z"][12:, 0] = np.arange(25) * 0.5 + 0.25           # z of layer 2 = 0.75
        A["hksat"][:] = 1e-2
        A["hk_l"][:] = 1e-2
        A["smp_l"][:] = -25000.0
        A["tsai"][:] = 0.25
        A["tlai"][:] = 0.75
        A["rootfr"][1] = 1.0
        c.phs_soil_root_conductance(A.layout(), A.host_struct(), 1)
        assert abs(A["root_conductance"][1, 0] - 0.999415208562283) <= 1e-13
        assert A["k_soil_root"][1, 0] > 0 and A["k_soil_root"][0, 0] == 0.0
    finally:
        c.close()


def test_error_paths(ctx):
    from ctsm_amd.lib import CTSM_E_ARG, Context, CtsmError
    A = PH.synthetic(16, seed=31)
    c = Context()
    try:
        with pytest.raises(CtsmError) as ei:                   # no parameters yet
            c.phs_soil_root_conductance(A.layout(), A.host_struct(), A.npatch)
        assert ei.value.code == CTSM_E_ARG
    finally:
        c.close()
    with pytest.raises(CtsmError) as ei:                       # a member the call needs is missing
        ctx.phs_root_sink(A.layout(), A.host_struct(omit=("vegwp",)), A.ncol)
    assert ei.value.code == CTSM_E_ARG
    with pytest.raises(CtsmError) as ei:                       # filter outside the allocation
        ctx.phs_root_sink(A.layout(), A.host_struct(), 1, np.array([A.ncol + 5], dtype=np.int32))
    assert ei.value.code == CTSM_E_ARG
    B = A.copy()
    B["itype"][0] = 70                                         # beyond the 17 types of the table
    with pytest.raises(CtsmError) as ei:
        ctx.phs_soil_root_conductance(B.layout(), B.host_struct(), B.npatch)
    assert ei.value.code == CTSM_E_ARG
    # the context still works after a reported error
    ref = A.copy()
    assert O.phs_root_sink(ref) == 0
    ctx.phs_root_sink(A.layout(), A.host_struct(), A.ncol)
    _same(A, ref, SINK)


STRESS = ("vegwp", "vegwp_pd", "bsun", "bsha", "qflx_tran_veg")


@pytest.mark.parametrize("ncol,seed", [(1, 51), (700, 52), (30000, 53)])
def test_calcstress_dropin(ctx, ncol, seed):
    print(f"case: calcstress drop-in, {ncol} columns, seed {seed}", flush=True)
    A = PH.synthetic_stress(ncol, seed=seed)
    ref = A.copy()
    assert O.phs_calcstress(ref, PH.default_params()) == 0
    ctx.phs_calcstress(A.layout(), A.host_struct(), A.npatch)
    _same(A, ref, STRESS)


def test_calcstress_dev_with_filter(ctx):
    print("case: calcstress device entry with a patch filter", flush=True)
    A = PH.synthetic_stress(5000, seed=54)
    rng = np.random.default_rng(55)
    filt = np.sort(rng.choice(np.arange(1, A.npatch + 1), A.npatch // 2, replace=False)).astype(np.int32)
    ref = A.copy()
    assert O.phs_calcstress(ref, PH.default_params(), filt) == 0
    D = PH.DevicePhs(ctx, A)
    fd = ctx.alloc(filt.nbytes)
    ctx.h2d(fd, filt)
    ctx.phs_calcstress(A.layout(), D.struct(), len(filt), fd, dev=True)
    out = D.fetch()
    _same(out, ref, STRESS)
    ctx.free(fd)
    D.free()


def test_calcstress_error_paths(ctx):
    from ctsm_amd.lib import CTSM_E_ARG, Context, CtsmError
    A = PH.synthetic_stress(8, seed=56)
    c = Context()
    try:
        p = PH.default_params()
        c.set_phs_params(PH.phs_params(*p._keep[:6]))           # no kmax
        with pytest.raises(CtsmError) as ei:
            c.phs_calcstress(A.layout(), A.host_struct(), A.npatch)
        assert ei.value.code == CTSM_E_ARG
    finally:
        c.close()
    with pytest.raises(CtsmError) as ei:
        ctx.phs_calcstress(A.layout(), A.host_struct(omit=("qaf",)), A.npatch)
    assert ei.value.code == CTSM_E_ARG
    B = A.copy()
    B["itype"][2] = 99
    with pytest.raises(CtsmError) as ei:
        ctx.phs_calcstress(B.layout(), B.host_struct(), B.npatch)
    assert ei.value.code == CTSM_E_ARG
    ref = A.copy()
    assert O.phs_calcstress(ref, PH.default_params()) == 0
    ctx.phs_calcstress(A.layout(), A.host_struct(), A.npatch)
    _same(A, ref, STRESS)
