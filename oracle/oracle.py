"""ctypes loader of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The product library never calls it.
"""
import ctypes as C
import os

import numpy as np

from ctsm_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        # CTSM_ORACLE_LIB points the checker at another build of the same
        # sources (the planted-mistake builds of tests/test_oracle_mutants.py)
        path = os.environ.get("CTSM_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.oracle_dgtsv.argtypes = [C.c_int, abi.f64p, abi.f64p, abi.f64p, abi.f64p]
        L.oracle_dgbsv.argtypes = [C.c_int, C.c_int, C.c_int, abi.f64p, C.c_int, abi.i32p, abi.f64p]
        sig = [C.POINTER(abi.Config), C.POINTER(abi.Layout), C.POINTER(abi.ColState)]
        L.oracle_soilwater.argtypes = sig + [C.c_int, abi.i32p, C.c_double, abi.i32p]
        L.oracle_soilwater_fractions.argtypes = sig + [C.c_int, abi.i32p]
        L.oracle_soiltemperature.argtypes = sig + [C.POINTER(abi.PatchState), C.c_int, abi.i32p,
                                                   C.c_double, abi.i32p]
        L.oracle_soil_step.argtypes = sig + [C.POINTER(abi.PatchState), C.c_int, abi.i32p,
                                             C.c_double, abi.i32p]
        L.oracle_config_default.argtypes = [C.POINTER(abi.Config)]
        _LIB = L
    return _LIB


def config(**kw):
    cfg = abi.Config()
    lib().oracle_config_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def _filt(filter_):
    if filter_ is None:
        return None, None
    f = np.ascontiguousarray(filter_, dtype=np.int32)
    return f, f.ctypes.data_as(abi.i32p)


def dgtsv(dl, d, du, b):
    dl, d, du, b = (np.array(x, dtype=np.float64) for x in (dl, d, du, b))
    info = lib().oracle_dgtsv(len(d), dl.ctypes.data_as(abi.f64p), d.ctypes.data_as(abi.f64p),
                              du.ctypes.data_as(abi.f64p), b.ctypes.data_as(abi.f64p))
    return b, info


def dgbsv(kl, ku, ab, b):
    """ab: (2*kl+ku+1, n) LAPACK band storage (rows kl..2kl+ku hold A)."""
    ab = np.asfortranarray(np.array(ab, dtype=np.float64))
    b = np.array(b, dtype=np.float64)
    n = ab.shape[1]
    ipiv = np.zeros(n, dtype=np.int32)
    info = lib().oracle_dgbsv(n, kl, ku, ab.ctypes.data_as(abi.f64p), ab.shape[0],
                              ipiv.ctypes.data_as(abi.i32p), b.ctypes.data_as(abi.f64p))
    return b, ipiv, info


def _run(fn, st, cfg, dtime, filter_, with_patch):
    cs, ps = st.structs()
    lay = st.layout()
    f, fp = _filt(filter_)
    n = st.dims.ncol if f is None else len(f)
    fail = C.c_int32(0)
    args = [C.byref(cfg), C.byref(lay), C.byref(cs)]
    if with_patch:
        args.append(C.byref(ps))
    args += [n, fp, C.c_double(dtime), C.byref(fail)]
    rc = fn(*args)
    return rc, fail.value


def soilwater(st, cfg=None, dtime=1800.0, filter_=None):
    return _run(lib().oracle_soilwater, st, cfg or config(), dtime, filter_, False)


def soiltemperature(st, cfg=None, dtime=1800.0, filter_=None):
    return _run(lib().oracle_soiltemperature, st, cfg or config(), dtime, filter_, True)


def soil_step(st, cfg=None, dtime=1800.0, filter_=None):
    return _run(lib().oracle_soil_step, st, cfg or config(), dtime, filter_, True)


def soilwater_fractions(st, cfg=None, filter_=None):
    cs, _ = st.structs()
    lay = st.layout()
    f, fp = _filt(filter_)
    n = st.dims.ncol if f is None else len(f)
    return lib().oracle_soilwater_fractions(C.byref(cfg or config()), C.byref(lay), C.byref(cs), n, fp)


_SYS_NAMES = ["hk", "smp", "dhkdw", "dsmpdw", "imped", "qin", "qout", "dqidw0", "dqidw1", "dqodw1",
              "dqodw2", "amx", "bmx", "cmx", "rmx", "dt_dz"]


def richards_system(st, col, dtsub, cfg=None):
    """The oracle's per-trial Richards system of 1-based column `col` at its
    current water content: dict of 1-D arrays over the column's layers."""
    L = lib()
    L.oracle_richards_system.argtypes = [C.POINTER(abi.Config), C.POINTER(abi.Layout), C.POINTER(abi.ColState),
                                         C.c_int, C.c_double, abi.f64p]
    cs, _ = st.structs()
    out = np.zeros(16 * 64)
    n = L.oracle_richards_system(C.byref(cfg or config()), C.byref(st.layout()), C.byref(cs), col, dtsub,
                                 out.ctypes.data_as(abi.f64p))
    assert n > 0
    return {k: out[i * n:(i + 1) * n].copy() for i, k in enumerate(_SYS_NAMES)}


def errsoi(st, tssbef, t_h2osfc_bef, t_grnd0, cfg=None, dtime=1800.0, filter_=None):
    """Per-patch soil energy balance error (SoilFluxesMod.F90:399-425) of a
    state after SoilTemperature, given the before-step temperatures."""
    L = lib()
    L.oracle_errsoi.argtypes = [C.POINTER(abi.Config), C.POINTER(abi.Layout), C.POINTER(abi.ColState),
                                C.POINTER(abi.PatchState), abi.f64p, abi.f64p, abi.f64p, C.c_int, abi.i32p,
                                C.c_double, abi.f64p]
    cs, ps = st.structs()
    f, fp = _filt(filter_)
    n = st.dims.ncol if f is None else len(f)
    tss = np.ascontiguousarray(tssbef, dtype=np.float64)
    thb = np.ascontiguousarray(t_h2osfc_bef, dtype=np.float64)
    tg0 = np.ascontiguousarray(t_grnd0, dtype=np.float64)
    err = np.zeros(st.dims.npatch)
    L.oracle_errsoi(C.byref(cfg or config()), C.byref(st.layout()), C.byref(cs), C.byref(ps),
                    tss.ctypes.data_as(abi.f64p), thb.ctypes.data_as(abi.f64p), tg0.ctypes.data_as(abi.f64p),
                    n, fp, C.c_double(dtime), err.ctypes.data_as(abi.f64p))
    return err


# ---- canopy path --------------------------------------------------------------

def _canopy_lib():
    L = lib()
    if not getattr(L, "_canopy_ready", False):
        L.oracle_canopyfluxes.argtypes = [C.POINTER(abi.Config), C.POINTER(abi.CanopyParams),
                                          C.POINTER(abi.PftParams), C.POINTER(abi.Layout),
                                          C.POINTER(abi.ColState), C.POINTER(abi.CanopyState), C.c_int, abi.i32p,
                                          C.c_double, abi.i32p]
        L.oracle_canopyfluxes_phs.argtypes = [C.POINTER(abi.Config), C.POINTER(abi.CanopyParams),
                                              C.POINTER(abi.PftParams), C.POINTER(abi.PhsParams),
                                              C.POINTER(abi.Layout), C.POINTER(abi.ColState),
                                              C.POINTER(abi.CanopyState), C.POINTER(abi.PhsState),
                                              C.POINTER(abi.PhsCanopyState), C.c_int, abi.i32p, C.c_double,
                                              abi.i32p]
        L.oracle_photosynthesis_phs_patch.argtypes = [C.POINTER(abi.Config), C.POINTER(abi.CanopyParams),
                                                      C.POINTER(abi.PftParams), C.POINTER(abi.PhsParams),
                                                      C.POINTER(abi.Layout), C.POINTER(abi.ColState),
                                                      C.POINTER(abi.CanopyState), C.POINTER(abi.PhsState),
                                                      C.POINTER(abi.PhsCanopyState), C.c_int] + [C.c_double] * 8
        L.oracle_root_sink.argtypes = [C.POINTER(abi.Config), C.POINTER(abi.Layout), C.POINTER(abi.ColState),
                                       C.POINTER(abi.CanopyState), C.c_int, abi.i32p]
        L.oracle_qsat_vec.argtypes = [C.c_int, abi.f64p, abi.f64p, abi.f64p, abi.f64p, abi.f64p]
        L.oracle_quadratic.argtypes = [C.c_double, C.c_double, C.c_double, abi.f64p, abi.f64p]
        L.oracle_frictionvelocity_vec.argtypes = [C.c_int] + [abi.f64p] * 11
        L.oracle_moninobukini.argtypes = [C.c_double] * 6 + [abi.f64p, abi.f64p]
        L.oracle_photosynthesis_patch.argtypes = [C.POINTER(abi.Config), C.POINTER(abi.CanopyParams),
                                                  C.POINTER(abi.PftParams), C.POINTER(abi.Layout),
                                                  C.POINTER(abi.ColState), C.POINTER(abi.CanopyState),
                                                  C.c_int, C.c_int] + [C.c_double] * 7
        for fn, na in (("oracle_ft", 2), ("oracle_fth", 4), ("oracle_fth25", 2)):
            getattr(L, fn).argtypes = [C.c_double] * na
            getattr(L, fn).restype = C.c_double
        L._canopy_ready = True
    return L


def canopy_config(**kw):
    """Oracle defaults plus the CLM6 canopy switches the synthetic workloads use."""
    base = dict(z0param_method=1, stomatalcond_method=2, soil_resis_method=1, use_biomass_heat_storage=1,
                use_undercanopy_stability=0, light_inhibit=1)
    base.update(kw)
    return config(**base)


def canopy_params_struct(d):
    s = abi.CanopyParams()
    for k in abi.CANOPY_PARAM_NAMES:
        setattr(s, k, float(d[k]))
    return s


class PftStruct:
    """abi.PftParams over numpy copies of a pft table (kept alive here)."""

    def __init__(self, table):
        self.arrays = {}
        self.s = abi.PftParams()
        n = None
        for name, t in abi.PFT_FIELDS:
            a = np.ascontiguousarray(table[name], dtype=np.int32 if t == "i32" else np.float64)
            n = len(a)
            self.arrays[name] = a
            setattr(self.s, name, a.ctypes.data_as(abi._CT[t]))
        self.s.npft = n


def canopyfluxes(st, cfg, prm, pft, filter_p=None, dtime=1800.0):
    L = _canopy_lib()
    cs, _ = st.structs()
    can = st.canopy_struct()
    f, fp = _filt(filter_p if filter_p is not None else np.arange(1, st.dims.npatch + 1))
    bad = C.c_int32(0)
    rc = L.oracle_canopyfluxes(C.byref(cfg), C.byref(canopy_params_struct(prm)), C.byref(pft.s),
                               C.byref(st.layout()), C.byref(cs), C.byref(can), len(f), fp, C.c_double(dtime),
                               C.byref(bad))
    return rc, bad.value


def canopyfluxes_phs(st, cphs, cfg, prm, pft, pp, filter_p=None, dtime=1800.0):
    """CanopyFluxes with plant hydraulic stress (oracle_canopyfluxes_phs) on
    the land state st and its CanopyPhs arrays cphs; cfg.use_hydrstress is
    not read (the PHS path is this call)."""
    L = _canopy_lib()
    cs, _ = st.structs()
    can = st.canopy_struct()
    phs, phc = cphs.host_structs(st)
    f, fp = _filt(filter_p if filter_p is not None else np.arange(1, st.dims.npatch + 1))
    bad = C.c_int32(0)
    rc = L.oracle_canopyfluxes_phs(C.byref(cfg), C.byref(canopy_params_struct(prm)), C.byref(pft.s), C.byref(pp),
                                   C.byref(st.layout()), C.byref(cs), C.byref(can), C.byref(phs), C.byref(phc),
                                   len(f), fp, C.c_double(dtime), C.byref(bad))
    return rc, bad.value


def photosynthesis_phs_patch(st, cphs, cfg, prm, pft, pp, p, esat_tv, eair, oair, cair, rb, dayl_factor, qsatl, qaf):
    """PhotosynthesisHydraulicStress for patch p (1-based) at the leaf
    temperature in st.can['t_veg'] (test access; the conductances must be in
    cphs['k_soil_root']); outputs are written into st and cphs."""
    L = _canopy_lib()
    cs, _ = st.structs()
    can = st.canopy_struct()
    phs, phc = cphs.host_structs(st)
    return L.oracle_photosynthesis_phs_patch(C.byref(cfg), C.byref(canopy_params_struct(prm)), C.byref(pft.s),
                                             C.byref(pp), C.byref(st.layout()), C.byref(cs), C.byref(can),
                                             C.byref(phs), C.byref(phc), int(p), esat_tv, eair, oair, cair, rb,
                                             dayl_factor, qsatl, qaf)


def phs_conductance_for(st, cphs, pp, cfg=None, filter_p=None):
    """The root-soil conductances of a land state's patches into cphs."""
    L = _phs_lib()
    phs, _ = cphs.host_structs(st)
    f, fp = _filt(filter_p if filter_p is not None else np.arange(1, st.dims.npatch + 1))
    return L.oracle_phs_soil_root_conductance(C.byref(cfg or config()), C.byref(st.layout()), C.byref(phs),
                                              C.byref(pp), len(f), fp)


def photosynthesis_patch(st, cfg, prm, pft, p, sun, esat_tv, eair, oair, cair, rb, btran, dayl_factor=1.0):
    """Photosynthesis for patch p (1-based) and one phase at the leaf
    temperature held in st.can['t_veg'] (test access to the oracle's
    routine; outputs are written into the state)."""
    L = _canopy_lib()
    cs, _ = st.structs()
    can = st.canopy_struct()
    return L.oracle_photosynthesis_patch(C.byref(cfg), C.byref(canopy_params_struct(prm)), C.byref(pft.s),
                                         C.byref(st.layout()), C.byref(cs), C.byref(can), int(p), int(sun),
                                         esat_tv, eair, oair, cair, rb, btran, dayl_factor)


def ft(tl, ha):
    return _canopy_lib().oracle_ft(tl, ha)


def fth(tl, hd, se, c):
    return _canopy_lib().oracle_fth(tl, hd, se, c)


def fth25(hd, se):
    return _canopy_lib().oracle_fth25(hd, se)


def root_sink(st, cfg, filter_c=None):
    L = _canopy_lib()
    cs, _ = st.structs()
    can = st.canopy_struct()
    f, fp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    L.oracle_root_sink(C.byref(cfg), C.byref(st.layout()), C.byref(cs), C.byref(can), len(f), fp)


def land_step(st, cfg, prm, pft, dtime=1800.0):
    """One land step with canopy, in clm_drv order: zero the ground-flux
    derivatives (BiogeophysPreFluxCalcsMod.F90:390-392), CanopyFluxes,
    SoilTemperature, SetSoilWaterFractions, the transpiration sink, SoilWater."""
    st.pch["cgrnd"][:] = 0.0
    st.can["cgrnds"][:] = 0.0
    st.can["cgrndl"][:] = 0.0
    rc, bad = canopyfluxes(st, cfg, prm, pft, dtime=dtime)
    if rc:
        return rc, bad
    rc, bad = soiltemperature(st, cfg, dtime)
    if rc:
        return rc, bad
    soilwater_fractions(st, cfg)
    root_sink(st, cfg)
    return soilwater(st, cfg, dtime)


# ---- SoilFluxes and the soil energy balance check --------------------------------

def _sf_lib():
    L = lib()
    if not getattr(L, "_sf_ready", False):
        P = C.POINTER
        L.oracle_save_before.argtypes = [P(abi.Config), P(abi.Layout), P(abi.ColState), P(abi.SoilFluxState),
                                         C.c_int, abi.i32p]
        L.oracle_soilfluxes.argtypes = [P(abi.Config), P(abi.Layout), P(abi.ColState), P(abi.PatchState),
                                        P(abi.SoilFluxState), C.c_int, abi.i32p, C.c_int, abi.i32p, C.c_double]
        L.oracle_soil_balance.argtypes = [P(abi.Layout), P(abi.SoilFluxState), C.c_int, abi.i32p, abi.f64p,
                                          P(C.c_int)]
        L.oracle_energy_balance.argtypes = [P(abi.Layout), P(abi.ColState), P(abi.PatchState), P(abi.CanopyState),
                                            P(abi.SoilFluxState), C.c_int, abi.i32p, abi.f64p, abi.f64p,
                                            abi.f64p, P(C.c_int), abi.f64p, P(C.c_int)]
        L._sf_ready = True
    return L


def save_before(st, cfg=None, filter_c=None):
    """t_ssbef / t_h2osfc_bef <- t_soisno / t_h2osfc (BiogeophysPreFluxCalcsMod.F90:302-331)."""
    L = _sf_lib()
    cs, _ = st.structs()
    sf = st.soilflux_struct()
    f, fp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    L.oracle_save_before(C.byref(cfg or canopy_config()), C.byref(st.layout()), C.byref(cs), C.byref(sf),
                         len(f), fp)


def soilfluxes(st, cfg=None, dtime=1800.0, filter_c=None, filter_p=None):
    """SoilFluxes (SoilFluxesMod.F90:38-518) on a state after SoilTemperature."""
    L = _sf_lib()
    cs, ps = st.structs()
    sf = st.soilflux_struct()
    fc, fcp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    fpp, fpp_p = _filt(filter_p if filter_p is not None else np.arange(1, st.dims.npatch + 1))
    return L.oracle_soilfluxes(C.byref(cfg or canopy_config()), C.byref(st.layout()), C.byref(cs), C.byref(ps),
                               C.byref(sf), len(fc), fcp, len(fpp), fpp_p, C.c_double(dtime))


def soil_balance(st, filter_c=None):
    """(max |errsoi_col|, column index) as BalanceCheckMod.F90:1101-1119 finds them."""
    L = _sf_lib()
    sf = st.soilflux_struct()
    f, fp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    m = C.c_double(0.0)
    idx = C.c_int(0)
    L.oracle_soil_balance(C.byref(st.layout()), C.byref(sf), len(f), fp, C.byref(m), C.byref(idx))
    return m.value, idx.value


def energy_balance(st, filter_p=None):
    """EnergyBalanceCheck's patch errors (BalanceCheckMod.F90:966-1003):
    (errseb, errlon per filter entry, (max |errseb|, patch), (max |errlon|, patch))."""
    L = _sf_lib()
    cs, ps = st.structs()
    f, fp = _filt(filter_p if filter_p is not None else np.arange(1, st.dims.npatch + 1))
    seb, lon = np.zeros(len(f)), np.zeros(len(f))
    ms, ml = C.c_double(0.0), C.c_double(0.0)
    i_s, i_l = C.c_int(0), C.c_int(0)
    L.oracle_energy_balance(C.byref(st.layout()), C.byref(cs), C.byref(ps), C.byref(st.canopy_struct()),
                            C.byref(st.soilflux_struct()), len(f), fp, seb.ctypes.data_as(abi.f64p),
                            lon.ctypes.data_as(abi.f64p), C.byref(ms), C.byref(i_s), C.byref(ml), C.byref(i_l))
    return seb, lon, (ms.value, i_s.value), (ml.value, i_l.value)


# ---- the column water mass and balance, the solar balance ------------------------

def _bal_lib():
    L = lib()
    if not getattr(L, "_bal_ready", False):
        P = C.POINTER
        L.oracle_water_mass.argtypes = [P(abi.Config), P(abi.Layout), P(abi.ColState), P(abi.CanopyState),
                                        P(abi.WaterBalanceState), C.c_int, abi.i32p, C.c_double, C.c_int, abi.f64p]
        L.oracle_water_balance.argtypes = [P(abi.Layout), P(abi.WaterBalanceState), C.c_int, abi.i32p, C.c_double,
                                           abi.f64p, P(C.c_int)]
        L.oracle_solar_balance.argtypes = [P(abi.Layout), P(abi.CanopyState), C.c_int, abi.i32p, abi.f64p, abi.f64p,
                                           abi.f64p, abi.f64p, abi.f64p, abi.f64p, P(C.c_int)]
        L._bal_ready = True
    return L


def water_mass(st, wb, cfg=None, filter_c=None, baseline=5000.0, subtract=False):
    """ComputeWaterMassNonLake (TotalWaterAndHeatMod.F90:92-141, 200-393): the
    column water mass (mm) per column of st (0 outside the filter); wb is a
    ctsm_amd.balance.WaterBalance."""
    L = _bal_lib()
    cs, _ = st.structs()
    out = np.zeros(st.dims.ncol)
    f, fp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    L.oracle_water_mass(C.byref(cfg or canopy_config()), C.byref(st.layout()), C.byref(cs),
                        C.byref(st.canopy_struct()), C.byref(wb.struct()), len(f), fp, C.c_double(baseline),
                        int(bool(subtract)), out.ctypes.data_as(abi.f64p))
    return out


def water_balance(st, wb, dtime=1800.0, filter_c=None):
    """errh2o per column into wb["errh2o"] (when present) and (max |errh2o|,
    column) as BalanceCheckMod.F90:588-620 finds them."""
    L = _bal_lib()
    f, fp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    m, idx = C.c_double(0.0), C.c_int(0)
    L.oracle_water_balance(C.byref(st.layout()), C.byref(wb.struct()), len(f), fp, C.c_double(dtime), C.byref(m),
                           C.byref(idx))
    return m.value, idx.value


def solar_balance(st, fsa, fsr, forc_solad_col, forc_solai, filter_p=None):
    """errsol per filter entry and (max |errsol|, patch) (BalanceCheckMod.F90:
    964-970, 1008-1012); forc_solad_col is (2, ncol), forc_solai (2, ngrid)."""
    L = _bal_lib()
    f, fp = _filt(filter_p if filter_p is not None else np.arange(1, st.dims.npatch + 1))
    err = np.zeros(len(f))
    m, idx = C.c_double(0.0), C.c_int(0)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (fsa, fsr, forc_solad_col, forc_solai)]
    L.oracle_solar_balance(C.byref(st.layout()), C.byref(st.canopy_struct()), len(f), fp,
                           *[a.ctypes.data_as(abi.f64p) for a in arrs], err.ctypes.data_as(abi.f64p), C.byref(m),
                           C.byref(idx))
    return err, (m.value, idx.value)


def _hy_lib():
    L = lib()
    if not getattr(L, "_hy_ready", False):
        P = C.POINTER
        L.oracle_surface_water.argtypes = [P(abi.Config), P(abi.Layout), P(abi.ColState), P(abi.SurfaceWaterState),
                                           P(abi.HydroParams), C.c_int, abi.i32p, C.c_double]
        L.oracle_patch2col3.argtypes = [P(abi.Layout), P(abi.ColState), abi.i32p, abi.f64p] + [abi.f64p] * 6 + [
            C.c_int, abi.i32p]
        L.oracle_water_table.argtypes = [P(abi.Config), P(abi.Layout), P(abi.ColState), P(abi.WaterTableState),
                                         C.c_int, abi.i32p, C.c_double]
        L.oracle_drainage.argtypes = [P(abi.Config), P(abi.Layout), P(abi.ColState), P(abi.DrainageState),
                                      P(abi.HydroParams), C.c_int, abi.i32p, C.c_double]
        L._hy_ready = True
    return L


def _phs_lib():
    L = lib()
    if not getattr(L, "_phs_ready", False):
        P = C.POINTER
        L.oracle_phs_plc.argtypes = [C.c_double] * 3
        L.oracle_phs_plc.restype = C.c_double
        L.oracle_phs_d1plc.argtypes = [C.c_double] * 3
        L.oracle_phs_d1plc.restype = C.c_double
        L.oracle_phs_soil_root_conductance.argtypes = [P(abi.Config), P(abi.Layout), P(abi.PhsState),
                                                       P(abi.PhsParams), C.c_int, abi.i32p]
        L.oracle_phs_root_sink.argtypes = [P(abi.Config), P(abi.Layout), P(abi.PhsState), C.c_int, abi.i32p]
        L.oracle_phs_calcstress.argtypes = [P(abi.Config), P(abi.Layout), P(abi.PhsStressState), P(abi.PhsParams),
                                            C.c_int, abi.i32p]
        L.oracle_phs_spac.argtypes = [P(abi.Config), P(abi.Layout), P(abi.PhsStressState), P(abi.PhsParams), C.c_int,
                                      abi.f64p, C.c_double, C.c_double, abi.f64p, abi.f64p]
        L._phs_ready = True
    return L


def phs_plc(x, psi50, ck):
    """plc, vegetation_weibull (PhotosynthesisMod.F90:5167-5195)."""
    return _phs_lib().oracle_phs_plc(x, psi50, ck)


def phs_d1plc(x, psi50, ck):
    """d1plc (PhotosynthesisMod.F90:5199-5228)."""
    return _phs_lib().oracle_phs_d1plc(x, psi50, ck)


def _phs_cfg(A):
    return abi.Config(nlevsno=A.dims["nlevsno"], nlevsoi=A.dims["nlevsoi"], nlevgrnd=A.dims["nlevgrnd"])


def phs_soil_root_conductance(A, prm, filter_p=None):
    """Root-soil conductances (PhotosynthesisMod.F90:3062-3114) on a
    ctsm_amd.phs.PhsArrays, in place; filter_p defaults to every patch."""
    f, fp = _filt(filter_p if filter_p is not None else np.arange(A.begp, A.begp + A.npatch))
    s = A.host_struct()
    return _phs_lib().oracle_phs_soil_root_conductance(C.byref(_phs_cfg(A)), C.byref(A.layout()), C.byref(s),
                                                       C.byref(prm), len(f), fp)


def phs_root_sink(A, filter_c=None):
    """Compute_EffecRootFrac_And_VertTranSink_HydStress
    (SoilWaterPlantSinkMod.F90:236-328) on a PhsArrays, in place."""
    f, fp = _filt(filter_c if filter_c is not None else np.arange(A.begc, A.begc + A.ncol))
    s = A.host_struct()
    return _phs_lib().oracle_phs_root_sink(C.byref(_phs_cfg(A)), C.byref(A.layout()), C.byref(s), len(f), fp)


def phs_calcstress(A, prm, filter_p=None):
    """calcstress (PhotosynthesisMod.F90:4490-4710) for every patch of the
    filter of a ctsm_amd.phs.StressArrays, in place."""
    f, fp = _filt(filter_p if filter_p is not None else np.arange(A.begp, A.begp + A.npatch))
    s = A.host_struct()
    return _phs_lib().oracle_phs_calcstress(C.byref(_phs_cfg(A)), C.byref(A.layout()), C.byref(s), C.byref(prm),
                                            len(f), fp)


def phs_spac(A, prm, p, x, q_sun, q_sha):
    """(f, invA, flag) of spacF / spacA (:4715-4976) for patch p at x."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.zeros(4)
    ia = np.zeros(16)
    s = A.host_struct()
    flag = _phs_lib().oracle_phs_spac(C.byref(_phs_cfg(A)), C.byref(A.layout()), C.byref(s), C.byref(prm), p,
                                      x.ctypes.data_as(abi.f64p), q_sun, q_sha, f.ctypes.data_as(abi.f64p),
                                      ia.ctypes.data_as(abi.f64p))
    return f, ia.reshape(4, 4), flag


def hydro_params(fff=0.5, pc=0.4, mu=0.14, h2osfcflag=1, aq_sp_yield_min=0.01, n_baseflow=1.0,
                 baseflow_scalar=5.5e-3, perched_baseflow_scalar=1.0e-5):
    return abi.HydroParams(fff=fff, pc=pc, mu=mu, aq_sp_yield_min=aq_sp_yield_min, n_baseflow=n_baseflow,
                           baseflow_scalar=baseflow_scalar, perched_baseflow_scalar=perched_baseflow_scalar,
                           h2osfcflag=h2osfcflag, pad=0)


def surface_water(st, hydro, cfg=None, dtime=1800.0, filter_c=None):
    """The surface water input chain (HydrologyNoDrainageMod.F90:306-336)."""
    L = _hy_lib()
    cs, _ = st.structs()
    sw = st.surfwater_struct()
    f, fp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    return L.oracle_surface_water(C.byref(cfg or canopy_config()), C.byref(st.layout()), C.byref(cs), C.byref(sw),
                                  C.byref(hydro), len(f), fp, C.c_double(dtime))


def patch2col3(st, filter_c=None, dew=False):
    """Column values of the three evaporation fluxes SetQflxInputs reads, or
    (dew=True) of the dew / sublimation fluxes RenewCondensation reads."""
    L = _hy_lib()
    cs, _ = st.structs()
    f, fp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    g, hy = st.can, st.hy
    if dew:
        names = ("qflx_liqdew_to_top_layer", "qflx_soliddew_to_top_layer", "qflx_solidevap_from_top_layer")
        arrs = [st.canopy_array("active"), st.canopy_array("wtcol")] + [st.sf[k] for k in names] + [
            hy[k] for k in names]
    else:
        arrs = [st.canopy_array("active"), st.canopy_array("wtcol"), st.sf["qflx_liqevap_from_top_layer"],
                st.canopy_array("qflx_ev_soil"), st.canopy_array("qflx_ev_h2osfc"),
                hy["qflx_liqevap_from_top_layer"], hy["qflx_ev_soil"], hy["qflx_ev_h2osfc"]]
    ptrs = [arrs[0].ctypes.data_as(abi.i32p)] + [a.ctypes.data_as(abi.f64p) for a in arrs[1:]]
    L.oracle_patch2col3(C.byref(st.layout()), C.byref(cs), *ptrs, len(f), fp)


def water_table(st, cfg=None, dtime=1800.0, filter_c=None):
    """PerchedWaterTable, ThetaBasedWaterTable and RenewCondensation
    (HydrologyNoDrainageMod.F90:355-373).  Returns 0, or the first column
    whose top-layer ice sublimation leaves it negative (the endrun)."""
    L = _hy_lib()
    cs, _ = st.structs()
    wt = st.watertable_struct()
    f, fp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    return L.oracle_water_table(C.byref(cfg or canopy_config()), C.byref(st.layout()), C.byref(cs), C.byref(wt),
                                len(f), fp, C.c_double(dtime))


def drainage(st, hydro, cfg=None, dtime=1800.0, filter_c=None):
    """PerchedLateralFlow, SubsurfaceLateralFlow and the runoff sum of
    HydrologyDrainage (HydrologyDrainageMod.F90:137-245, no aquifer layer)."""
    L = _hy_lib()
    cs, _ = st.structs()
    dr = st.drainage_struct()
    f, fp = _filt(filter_c if filter_c is not None else np.arange(1, st.dims.ncol + 1))
    return L.oracle_drainage(C.byref(cfg or canopy_config()), C.byref(st.layout()), C.byref(cs), C.byref(dr),
                             C.byref(hydro), len(f), fp, C.c_double(dtime))


def land_step_full(st, cfg, prm, pft, dtime=1800.0, hydro=None, water_table_after=False, drainage_after=False):
    """land_step with the reference's bookkeeping: t_ssbef / t_h2osfc_bef
    saved first, SoilFluxes between SoilTemperature and the hydrology
    (clm_driver.F90:921), given hydro the column evaporation fluxes and the
    surface water input chain before SoilWater, and with water_table_after
    the column dew fluxes, the water tables and RenewCondensation after it."""
    save_before(st, cfg)
    st.pch["cgrnd"][:] = 0.0
    st.can["cgrnds"][:] = 0.0
    st.can["cgrndl"][:] = 0.0
    rc, bad = canopyfluxes(st, cfg, prm, pft, dtime=dtime)
    if rc:
        return rc, bad
    rc, bad = soiltemperature(st, cfg, dtime)
    if rc:
        return rc, bad
    soilwater_fractions(st, cfg)
    soilfluxes(st, cfg, dtime)
    root_sink(st, cfg)
    if hydro is not None:
        patch2col3(st)
        surface_water(st, hydro, cfg, dtime)
    rc, bad = soilwater(st, cfg, dtime)
    if rc:
        return rc, bad
    if water_table_after:
        patch2col3(st, dew=True)
        c = water_table(st, cfg, dtime)
        if c:
            return 1, c
    if drainage_after:
        drainage(st, hydro, cfg, dtime)
    return 0, None


def qsat(T, p):
    L = _canopy_lib()
    T = np.ascontiguousarray(T, dtype=np.float64)
    p = np.ascontiguousarray(np.broadcast_to(p, T.shape), dtype=np.float64)
    qs, es, qsdT = (np.zeros_like(T) for _ in range(3))
    L.oracle_qsat_vec(len(T), *(a.ctypes.data_as(abi.f64p) for a in (T, p, qs, es, qsdT)))
    return qs, es, qsdT


def quadratic(a, b, c):
    L = _canopy_lib()
    r1, r2 = np.zeros(1), np.zeros(1)
    rc = L.oracle_quadratic(a, b, c, r1.ctypes.data_as(abi.f64p), r2.ctypes.data_as(abi.f64p))
    return rc, r1[0], r2[0]


_FV_NAMES = ["ustar", "temp1", "temp2", "temp12m", "temp22m", "fm", "u10", "u10_clm", "va", "fv", "vds"]


def frictionvelocity(hu, ht, hq, displa, z0m, z0h, z0q, obu, ur, um):
    L = _canopy_lib()
    args = [np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64) for x in
            np.broadcast_arrays(hu, ht, hq, displa, z0m, z0h, z0q, obu, ur, um)]
    n = len(args[0])
    out = np.zeros(11 * n)
    L.oracle_frictionvelocity_vec(n, *(a.ctypes.data_as(abi.f64p) for a in args), out.ctypes.data_as(abi.f64p))
    out = out.reshape(n, 11)
    return {k: out[:, i] for i, k in enumerate(_FV_NAMES)}


def moninobukini(ur, thv, dthv, zldis, z0m, zetamax):
    L = _canopy_lib()
    um, obu = np.zeros(1), np.zeros(1)
    L.oracle_moninobukini(ur, thv, dthv, zldis, z0m, zetamax, um.ctypes.data_as(abi.f64p),
                          obu.ctypes.data_as(abi.f64p))
    return um[0], obu[0]
