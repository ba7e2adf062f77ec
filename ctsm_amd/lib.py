"""Loader of libctsm_mi355x.so and thin Python mirrors of its entry points.

The Python side mirrors the reference's call sites: soilwater() stands for
SoilWater (SoilWaterMovementMod.F90:240), soiltemperature() for
SoilTemperature (SoilTemperatureMod.F90:92), both taking the state and a
column filter and mutating the state in place, and raising CtsmError where
the reference would call endrun.  There is no CPU fallback: if the HIP
library is missing or no GPU is visible, every call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libctsm_mi355x.so")
_LIB = None

(CTSM_OK, CTSM_E_ARG, CTSM_E_CONFIG, CTSM_E_HIP, CTSM_E_SOLVER, CTSM_E_NODEVICE, CTSM_E_ENDRUN, CTSM_E_IO,
 CTSM_E_SCRATCH) = range(9)
_NAMES = {1: "CTSM_E_ARG", 2: "CTSM_E_CONFIG", 3: "CTSM_E_HIP", 4: "CTSM_E_SOLVER", 5: "CTSM_E_NODEVICE",
          6: "CTSM_E_ENDRUN", 7: "CTSM_E_IO", 8: "CTSM_E_SCRATCH"}
NOALLOC_SCRATCH_BYTES = 128 << 20      # CTSM_NOALLOC_SCRATCH_BYTES

# every symbol declared in include/ctsm_gpu.h
EXPORTS = [
    "ctsm_config_default", "ctsm_create", "ctsm_destroy", "ctsm_last_error", "ctsm_set_tuning", "ctsm_get_tuning",
    "ctsm_soilwater", "ctsm_soiltemperature", "ctsm_soilwater_fractions",
    "ctsm_soilwater_dev", "ctsm_soiltemperature_dev", "ctsm_soilwater_fractions_dev",
    "ctsm_soil_step_fused_dev", "ctsm_dev_alloc", "ctsm_dev_free", "ctsm_memcpy", "ctsm_dev_copy",
    "ctsm_event_record", "ctsm_event_elapsed", "ctsm_event_destroy", "ctsm_sync", "ctsm_stream",
    "ctsm_set_canopy_params", "ctsm_canopyfluxes", "ctsm_root_sink", "ctsm_canopyfluxes_dev",
    "ctsm_canopyfluxes_phs_dev", "ctsm_canopyfluxes_phs",
    "ctsm_root_sink_dev", "ctsm_land_step_dev", "ctsm_memcpy2d",
    "ctsm_graph_capture", "ctsm_graph_launch", "ctsm_graph_destroy",
    "ctsm_soilfluxes", "ctsm_soil_balance_check", "ctsm_soilfluxes_dev", "ctsm_soil_balance_check_dev",
    "ctsm_land_step_full_dev", "ctsm_soil_steps_fused_dev", "ctsm_surface_water", "ctsm_surface_water_dev",
    "ctsm_water_table", "ctsm_water_table_dev",
    "ctsm_hist_update_dev", "ctsm_hist_normalize_dev", "ctsm_hist_zero_dev",
    "ctsm_checkpoint_write", "ctsm_checkpoint_read", "ctsm_drainage", "ctsm_drainage_dev",
    "ctsm_graph_capture_full", "ctsm_set_phs_params", "ctsm_phs_soil_root_conductance",
    "ctsm_phs_soil_root_conductance_dev", "ctsm_phs_root_sink", "ctsm_phs_root_sink_dev",
    "ctsm_phs_calcstress", "ctsm_phs_calcstress_dev", "ctsm_canopyfluxes_dev_noalloc",
    "ctsm_energy_balance_check_dev", "ctsm_set_pft_rows", "ctsm_serve_create", "ctsm_serve_submit", "ctsm_serve_progress", "ctsm_serve_poll",
    "ctsm_serve_start", "ctsm_serve_release", "ctsm_serve_resume_member", "ctsm_serve_stats", "ctsm_serve_destroy",
    "ctsm_solar_balance_check_dev", "ctsm_water_mass_dev", "ctsm_water_balance_check_dev",
]
SERVE_QUEUED, SERVE_ADMITTED, SERVE_RUNNING, SERVE_DONE = range(4)


class CtsmError(RuntimeError):
    """Raised where the reference calls endrun (e.g. LAPACK info != 0)."""

    def __init__(self, code, index=0, where=""):
        self.code, self.index = code, index
        super().__init__(f"{where}: {_NAMES.get(code, code)} (first failing column / patch {index})")


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.ctsm_config_default.argtypes = [P(abi.Config)]
    L.ctsm_config_default.restype = None
    L.ctsm_create.argtypes = [P(abi.Config), C.c_int, P(C.c_void_p)]
    L.ctsm_destroy.argtypes = [C.c_void_p]
    L.ctsm_last_error.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_int32)]
    L.ctsm_set_tuning.argtypes = [C.c_void_p, P(abi.Tuning)]
    L.ctsm_get_tuning.argtypes = [C.c_void_p, P(abi.Tuning)]
    col = [C.c_void_p, P(abi.Layout), P(abi.ColState)]
    L.ctsm_soilwater.argtypes = col + [C.c_int32, abi.i32p, C.c_double]
    L.ctsm_soilwater_fractions.argtypes = col + [C.c_int32, abi.i32p]
    L.ctsm_soiltemperature.argtypes = col + [P(abi.PatchState), C.c_int32, abi.i32p, C.c_double]
    L.ctsm_soilwater_dev.argtypes = col + [C.c_int32, C.c_void_p, C.c_double, C.c_int32]
    L.ctsm_soilwater_fractions_dev.argtypes = col + [C.c_int32, C.c_void_p]
    L.ctsm_soiltemperature_dev.argtypes = col + [P(abi.PatchState), C.c_int32, C.c_void_p, C.c_double, C.c_int32]
    L.ctsm_soil_step_fused_dev.argtypes = col + [P(abi.PatchState), C.c_int32, C.c_void_p, C.c_double]
    L.ctsm_soil_steps_fused_dev.argtypes = col + [P(abi.PatchState), C.c_int32, C.c_void_p, C.c_double, C.c_int32]
    L.ctsm_graph_capture.argtypes = col + [P(abi.PatchState), P(abi.CanopyState), C.c_int32, C.c_void_p, C.c_int32,
                                           C.c_void_p, C.c_double, C.c_int32, P(C.c_void_p)]
    sfa = col + [P(abi.PatchState), P(abi.SoilFluxState), C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_double]
    L.ctsm_soilfluxes.argtypes = sfa
    L.ctsm_soilfluxes_dev.argtypes = sfa
    bal = [C.c_void_p, P(abi.Layout), P(abi.SoilFluxState), C.c_int32, C.c_void_p, P(C.c_double), P(C.c_int32)]
    L.ctsm_soil_balance_check.argtypes = bal
    L.ctsm_soil_balance_check_dev.argtypes = bal
    L.ctsm_land_step_full_dev.argtypes = col + [P(abi.PatchState), P(abi.CanopyState), P(abi.SoilFluxState),
                                                P(abi.SurfaceWaterState), P(abi.HydroParams),
                                                P(abi.WaterTableState), P(abi.DrainageState),
                                                C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_double]
    swa = col + [P(abi.SurfaceWaterState), P(abi.HydroParams), C.c_int32, C.c_void_p, C.c_double]
    L.ctsm_surface_water.argtypes = swa
    L.ctsm_surface_water_dev.argtypes = swa
    wta = col + [P(abi.WaterTableState), C.c_int32, C.c_void_p, C.c_double]
    L.ctsm_water_table.argtypes = wta
    L.ctsm_water_table_dev.argtypes = wta
    dra = col + [P(abi.DrainageState), P(abi.HydroParams), C.c_int32, C.c_void_p, C.c_double]
    L.ctsm_drainage.argtypes = dra
    L.ctsm_drainage_dev.argtypes = dra
    L.ctsm_set_phs_params.argtypes = [C.c_void_p, P(abi.PhsParams)]
    for fn in ("ctsm_phs_soil_root_conductance", "ctsm_phs_soil_root_conductance_dev", "ctsm_phs_root_sink",
               "ctsm_phs_root_sink_dev"):
        getattr(L, fn).argtypes = [C.c_void_p, P(abi.Layout), P(abi.PhsState), C.c_int32, C.c_void_p]
    for fn in ("ctsm_phs_calcstress", "ctsm_phs_calcstress_dev"):
        getattr(L, fn).argtypes = [C.c_void_p, P(abi.Layout), P(abi.PhsStressState), C.c_int32, C.c_void_p]
    for fn in ("ctsm_hist_update_dev", "ctsm_hist_normalize_dev", "ctsm_hist_zero_dev"):
        getattr(L, fn).argtypes = [C.c_void_p, P(abi.HistField), C.c_int32]
    for fn in ("ctsm_checkpoint_write", "ctsm_checkpoint_read"):
        getattr(L, fn).argtypes = [C.c_void_p, C.c_char_p, P(abi.CkptField), C.c_int32]
    L.ctsm_graph_capture_full.argtypes = col + [P(abi.PatchState), P(abi.CanopyState), P(abi.SoilFluxState),
                                                P(abi.SurfaceWaterState), P(abi.HydroParams),
                                                P(abi.WaterTableState), P(abi.DrainageState),
                                                C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_double, C.c_int32,
                                                P(C.c_void_p)]
    L.ctsm_graph_launch.argtypes = [C.c_void_p, C.c_void_p]
    L.ctsm_graph_destroy.argtypes = [C.c_void_p]
    L.ctsm_sync.argtypes = [C.c_void_p]
    L.ctsm_stream.argtypes = [C.c_void_p]
    L.ctsm_stream.restype = C.c_void_p
    L.ctsm_dev_alloc.argtypes = [C.c_void_p, C.c_uint64, P(C.c_void_p)]
    L.ctsm_dev_free.argtypes = [C.c_void_p, C.c_void_p]
    L.ctsm_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32]
    L.ctsm_dev_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
    L.ctsm_memcpy2d.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64,
                                C.c_int32]
    L.ctsm_event_record.argtypes = [C.c_void_p, P(C.c_void_p)]
    L.ctsm_event_elapsed.argtypes = [C.c_void_p, C.c_void_p, P(C.c_float)]
    L.ctsm_event_destroy.argtypes = [C.c_void_p]
    L.ctsm_set_canopy_params.argtypes = [C.c_void_p, P(abi.CanopyParams), P(abi.PftParams)]
    can = col + [P(abi.CanopyState)]
    L.ctsm_canopyfluxes.argtypes = can + [C.c_int32, abi.i32p, C.c_double]
    L.ctsm_root_sink.argtypes = can + [C.c_int32, abi.i32p]
    L.ctsm_canopyfluxes_dev.argtypes = can + [C.c_int32, C.c_void_p, C.c_double]
    L.ctsm_canopyfluxes_dev_noalloc.argtypes = can + [C.c_int32, C.c_void_p, C.c_double]
    L.ctsm_canopyfluxes_phs_dev.argtypes = can + [P(abi.PhsState), P(abi.PhsCanopyState), C.c_int32, C.c_void_p,
                                                  C.c_double]
    L.ctsm_canopyfluxes_phs.argtypes = can + [P(abi.PhsState), P(abi.PhsCanopyState), C.c_int32, abi.i32p, C.c_double]
    L.ctsm_set_pft_rows.argtypes = [C.c_void_p, C.c_int32, P(abi.PftParams)]
    L.ctsm_energy_balance_check_dev.argtypes = [C.c_void_p, P(abi.Layout), P(abi.ColState), P(abi.PatchState),
                                                P(abi.CanopyState), P(abi.SoilFluxState), C.c_int32, C.c_void_p,
                                                C.c_void_p, C.c_void_p, P(C.c_double), P(C.c_int32),
                                                P(C.c_double), P(C.c_int32)]
    L.ctsm_solar_balance_check_dev.argtypes = [C.c_void_p, P(abi.Layout), P(abi.CanopyState), C.c_int32, C.c_void_p,
                                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                               P(C.c_double), P(C.c_int32)]
    L.ctsm_water_mass_dev.argtypes = [C.c_void_p, P(abi.Layout), P(abi.ColState), P(abi.CanopyState),
                                      P(abi.WaterBalanceState), C.c_int32, C.c_void_p, C.c_double, C.c_int32,
                                      C.c_void_p]
    L.ctsm_water_balance_check_dev.argtypes = [C.c_void_p, P(abi.Layout), P(abi.WaterBalanceState), C.c_int32,
                                               C.c_void_p, C.c_double, P(C.c_double), P(C.c_int32)]
    L.ctsm_serve_create.argtypes = [C.c_void_p, P(abi.Layout), P(abi.ColState), P(abi.PatchState),
                                    P(abi.CanopyState), P(abi.SoilFluxState), P(abi.SurfaceWaterState),
                                    P(abi.HydroParams), P(abi.WaterTableState), P(abi.DrainageState), C.c_int32,
                                    C.c_int32, C.c_double, P(C.c_void_p)]
    L.ctsm_serve_submit.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(C.c_int64)]
    L.ctsm_serve_progress.argtypes = [C.c_void_p, C.c_int32]
    L.ctsm_serve_poll.argtypes = [C.c_void_p, C.c_int64] + [P(C.c_int32)] * 5
    L.ctsm_serve_start.argtypes = [C.c_void_p, C.c_int64]
    L.ctsm_serve_release.argtypes = [C.c_void_p, C.c_int64]
    L.ctsm_serve_resume_member.argtypes = [C.c_void_p] + [C.c_int32] * 7 + [P(C.c_int64)]
    L.ctsm_serve_stats.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_double), P(C.c_int32), P(C.c_double),
                                   P(C.c_int32), P(C.c_double), P(C.c_int32), P(C.c_int32)]
    L.ctsm_serve_destroy.argtypes = [C.c_void_p]
    L.ctsm_root_sink_dev.argtypes = can + [C.c_int32, C.c_void_p]
    L.ctsm_land_step_dev.argtypes = col + [P(abi.PatchState), P(abi.CanopyState), C.c_int32, C.c_void_p,
                                           C.c_int32, C.c_void_p, C.c_double]
    _LIB = L
    return L


def default_config(**kw):
    cfg = abi.Config()
    load().ctsm_config_default(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


class Context:
    """One library context (HIP stream + workspace) on one device."""

    def __init__(self, cfg=None, device=0):
        self.L = load()
        self.cfg = cfg if cfg is not None else default_config()
        h = C.c_void_p()
        rc = self.L.ctsm_create(C.byref(self.cfg), device, C.byref(h))
        if rc != CTSM_OK:
            raise CtsmError(rc, 0, "ctsm_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.ctsm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, where):
        if rc != CTSM_OK:
            code, idx = C.c_int32(), C.c_int32()
            self.L.ctsm_last_error(self.h, C.byref(code), C.byref(idx))
            raise CtsmError(rc, idx.value, where)

    def set_tuning(self, **kw):
        t = abi.Tuning()
        self.L.ctsm_get_tuning(self.h, C.byref(t))
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._check(self.L.ctsm_set_tuning(self.h, C.byref(t)), "ctsm_set_tuning")

    def get_tuning(self):
        t = abi.Tuning()
        self.L.ctsm_get_tuning(self.h, C.byref(t))
        return {k: getattr(t, k) for k, _ in abi.Tuning._fields_}

    @property
    def stream(self):
        return self.L.ctsm_stream(self.h)

    # ---- drop-in (host numpy state) ------------------------------------
    def soilwater(self, st, filter_=None, dtime=1800.0):
        cs, _ = st.structs()
        f, fp, n = _filter(filter_, st.dims.ncol)
        self._check(self.L.ctsm_soilwater(self.h, C.byref(st.layout()), C.byref(cs), n, fp, dtime), "SoilWater")

    def soiltemperature(self, st, filter_=None, dtime=1800.0):
        cs, ps = st.structs()
        f, fp, n = _filter(filter_, st.dims.ncol)
        self._check(self.L.ctsm_soiltemperature(self.h, C.byref(st.layout()), C.byref(cs), C.byref(ps), n, fp,
                                                dtime), "SoilTemperature")

    def soilwater_fractions(self, st, filter_=None):
        cs, _ = st.structs()
        f, fp, n = _filter(filter_, st.dims.ncol)
        self._check(self.L.ctsm_soilwater_fractions(self.h, C.byref(st.layout()), C.byref(cs), n, fp),
                    "SetSoilWaterFractions")

    # ---- canopy ------------------------------------------------------------
    def set_canopy_params(self, params, pft_table):
        """params: dict of abi.CANOPY_PARAM_NAMES; pft_table: dict of abi.PFT_FIELDS arrays."""
        prm = abi.CanopyParams()
        for k in abi.CANOPY_PARAM_NAMES:
            setattr(prm, k, float(params[k]))
        keep = []
        pft = abi.PftParams()
        for name, t in abi.PFT_FIELDS:
            a = np.ascontiguousarray(pft_table[name], dtype=np.int32 if t == "i32" else np.float64)
            keep.append(a)
            setattr(pft, name, a.ctypes.data_as(abi._CT[t]))
            pft.npft = len(a)
        self._check(self.L.ctsm_set_canopy_params(self.h, C.byref(prm), C.byref(pft)), "ctsm_set_canopy_params")

    def energy_balance_check_dev(self, ds, filterp_dev=None, n=None, errseb_dev=None, errlon_dev=None):
        """EnergyBalanceCheck's patch errors on a DeviceState after a full land
        step: ((max |errseb|, patch), (max |errlon|, patch)); per-patch values
        go to errseb_dev / errlon_dev when given."""
        cs, ps = ds.structs()
        n = ds.dims.npatch if n is None else n
        ms, ml = C.c_double(), C.c_double()
        i_s, i_l = C.c_int32(), C.c_int32()
        fp = C.c_void_p(filterp_dev) if filterp_dev is not None else None
        self._check(self.L.ctsm_energy_balance_check_dev(
            self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ps), C.byref(ds.canopy_struct()),
            C.byref(ds.soilflux_struct()), n, fp, C.c_void_p(errseb_dev) if errseb_dev else None,
            C.c_void_p(errlon_dev) if errlon_dev else None, C.byref(ms), C.byref(i_s), C.byref(ml), C.byref(i_l)),
            "energy balance check")
        return (ms.value, i_s.value), (ml.value, i_l.value)

    def solar_balance_check_dev(self, ds, fsa_dev, fsr_dev, solad_dev, solai_dev, filterp_dev=None, n=None,
                                errsol_dev=None):
        """errsol of EnergyBalanceCheck over device arrays fsa / fsr (p),
        forc_solad_col (c, 2) and forc_solai (g, 2): (max |errsol|, patch)."""
        n = ds.dims.npatch if n is None else n
        m, i = C.c_double(), C.c_int32()
        vp = lambda x: C.c_void_p(x) if x else None
        self._check(self.L.ctsm_solar_balance_check_dev(
            self.h, C.byref(ds.layout()), C.byref(ds.canopy_struct()), n, vp(filterp_dev), vp(fsa_dev), vp(fsr_dev),
            vp(solad_dev), vp(solai_dev), vp(errsol_dev), C.byref(m), C.byref(i)), "solar balance check")
        return m.value, i.value

    def water_mass_dev(self, ds, dwb, out_dev, filter_dev=None, n=None, baseline=5000.0, subtract=False):
        """ComputeWaterMassNonLake on device: the column water mass of each
        filter column into the device (c) array out_dev (enqueued)."""
        cs, _ = ds.structs()
        n = ds.dims.ncol if n is None else n
        self._check(self.L.ctsm_water_mass_dev(
            self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ds.canopy_struct()), C.byref(dwb.struct()), n,
            C.c_void_p(filter_dev) if filter_dev else None, baseline, int(bool(subtract)), C.c_void_p(out_dev)),
            "water mass")

    def water_balance_check_dev(self, ds, dwb, filter_dev=None, n=None, dtime=1800.0):
        """errh2o per column into the DeviceWaterBalance's errh2o (when it has
        one): (max |errh2o|, column)."""
        n = ds.dims.ncol if n is None else n
        m, i = C.c_double(), C.c_int32()
        self._check(self.L.ctsm_water_balance_check_dev(
            self.h, C.byref(ds.layout()), C.byref(dwb.struct()), n, C.c_void_p(filter_dev) if filter_dev else None,
            dtime, C.byref(m), C.byref(i)), "water balance check")
        return m.value, i.value

    def set_pft_rows(self, first_row, pft_table):
        """Rows first_row.. of the context's PFT table from pft_table (a
        perturbed member's own PFT parameters)."""
        keep = []
        pft = abi.PftParams()
        for name, t in abi.PFT_FIELDS:
            a = np.ascontiguousarray(pft_table[name], dtype=np.int32 if t == "i32" else np.float64)
            keep.append(a)
            setattr(pft, name, a.ctypes.data_as(abi._CT[t]))
            pft.npft = len(a)
        self._check(self.L.ctsm_set_pft_rows(self.h, int(first_row), C.byref(pft)), "ctsm_set_pft_rows")

    # ---- the ensemble server (continuous batching in a resident pool) -------
    def serve_create(self, ds, structs, hydro, water_table, drainage, pft_first, max_steps_per_tick, dtime=1800.0):
        cs, ps, can, sf = structs
        sw = ds.surfwater_struct()
        wt = ds.watertable_struct() if water_table else None
        dr = ds.drainage_struct() if drainage else None
        hp = hydro
        out = C.c_void_p()
        self._keep_serve = (cs, ps, can, sf, sw, wt, dr, hp)
        self._check(self.L.ctsm_serve_create(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ps), C.byref(can),
                                             C.byref(sf), C.byref(sw), C.byref(hp),
                                             C.byref(wt) if wt is not None else None,
                                             C.byref(dr) if dr is not None else None, int(pft_first),
                                             int(max_steps_per_tick), dtime, C.byref(out)), "ctsm_serve_create")
        return out.value

    def serve_submit(self, srv, ncol, npatch, npft, nsteps):
        t = C.c_int64()
        self._check(self.L.ctsm_serve_submit(srv, ncol, npatch, npft, nsteps, C.byref(t)), "ctsm_serve_submit")
        return t.value

    def serve_progress(self, srv, wait=False, quiesce=False):
        """One scheduler turn; quiesce waits for every tick in flight and
        admits and enqueues nothing."""
        mode = 2 if quiesce else (1 if wait else 0)
        self._check(self.L.ctsm_serve_progress(srv, mode), "ctsm_serve_progress")

    def serve_poll(self, srv, ticket):
        """(state, col0, pat0, pft0, steps left) of a ticket."""
        st, c0, p0, f0, lf = (C.c_int32() for _ in range(5))
        self._check(self.L.ctsm_serve_poll(srv, ticket, C.byref(st), C.byref(c0), C.byref(p0), C.byref(f0),
                                           C.byref(lf)), "ctsm_serve_poll")
        return st.value, c0.value, p0.value, f0.value, lf.value

    def serve_start(self, srv, ticket):
        self._check(self.L.ctsm_serve_start(srv, ticket), "ctsm_serve_start")

    def serve_release(self, srv, ticket):
        self._check(self.L.ctsm_serve_release(srv, ticket), "ctsm_serve_release")

    def serve_resume_member(self, srv, ncol, npatch, npft, left, col0, pat0, pft0):
        t = C.c_int64()
        self._check(self.L.ctsm_serve_resume_member(srv, ncol, npatch, npft, left, col0, pat0, pft0, C.byref(t)),
                    "ctsm_serve_resume_member")
        return t.value

    def serve_stats(self, srv):
        """(column-steps, worst |errsoi|, its pool column, ticks in flight,
        worst |errseb|, its pool patch, worst |errlon|, its pool patch)."""
        cs, w, wc, inf = C.c_int64(), C.c_double(), C.c_int32(), C.c_int32()
        seb, sp, lon, lp = C.c_double(), C.c_int32(), C.c_double(), C.c_int32()
        self._check(self.L.ctsm_serve_stats(srv, C.byref(cs), C.byref(w), C.byref(wc), C.byref(seb), C.byref(sp),
                                            C.byref(lon), C.byref(lp), C.byref(inf)), "ctsm_serve_stats")
        return cs.value, w.value, wc.value, inf.value, seb.value, sp.value, lon.value, lp.value

    def serve_destroy(self, srv):
        self.L.ctsm_serve_destroy(srv)

    def canopyfluxes(self, st, filter_p=None, dtime=1800.0):
        cs, _ = st.structs()
        can = st.canopy_struct()
        f, fp, n = _filter(filter_p, st.dims.npatch)
        self._check(self.L.ctsm_canopyfluxes(self.h, C.byref(st.layout()), C.byref(cs), C.byref(can), n, fp, dtime),
                    "CanopyFluxes")

    def canopyfluxes_phs(self, st, cphs, filter_p=None, dtime=1800.0):
        """Drop-in CanopyFluxes with plant hydraulic stress on the host state st
        and its phs.CanopyPhs arrays (ctsm_canopyfluxes_phs)."""
        cs, _ = st.structs()
        can = st.canopy_struct()
        phs, phc = cphs.host_structs(st)
        f, fp, n = _filter(filter_p, st.dims.npatch)
        self._check(self.L.ctsm_canopyfluxes_phs(self.h, C.byref(st.layout()), C.byref(cs), C.byref(can),
                                                 C.byref(phs), C.byref(phc), n, fp, dtime), "CanopyFluxes (PHS)")

    def root_sink(self, st, filter_c=None):
        cs, _ = st.structs()
        can = st.canopy_struct()
        f, fp, n = _filter(filter_c, st.dims.ncol)
        self._check(self.L.ctsm_root_sink(self.h, C.byref(st.layout()), C.byref(cs), C.byref(can), n, fp),
                    "root sink")

    def canopyfluxes_dev(self, ds, filterp_dev=None, n=None, dtime=1800.0, structs=None):
        if structs is None:
            cs, _ = ds.structs()
            can = ds.canopy_struct()
        else:
            cs, _, can = structs
        n = ds.dims.npatch if n is None else n
        fp = C.c_void_p(filterp_dev) if filterp_dev is not None else None
        self._check(self.L.ctsm_canopyfluxes_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(can), n, fp,
                                                 dtime), "CanopyFluxes")

    def canopyfluxes_phs_dev(self, ds, dphs, filterp_dev=None, n=None, dtime=1800.0):
        """CanopyFluxes with plant hydraulic stress on device state ds and its
        phs.DeviceCanopyPhs arrays (the context's config has use_hydrstress = 1)."""
        cs, _ = ds.structs()
        can = ds.canopy_struct()
        phs, phc = dphs.structs()
        n = ds.dims.npatch if n is None else n
        fp = C.c_void_p(filterp_dev) if filterp_dev is not None else None
        self._check(self.L.ctsm_canopyfluxes_phs_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(can),
                                                     C.byref(phs), C.byref(phc), n, fp, dtime), "CanopyFluxes (PHS)")

    def canopyfluxes_dev_noalloc(self, ds, filterp_dev=None, n=None, dtime=1800.0, structs=None):
        """canopyfluxes_dev through the entry that allocates nothing (its
        scratch is the library's fixed per-device array)."""
        if structs is None:
            cs, _ = ds.structs()
            can = ds.canopy_struct()
        else:
            cs, _, can = structs
        n = ds.dims.npatch if n is None else n
        fp = C.c_void_p(filterp_dev) if filterp_dev is not None else None
        self._check(self.L.ctsm_canopyfluxes_dev_noalloc(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(can), n,
                                                         fp, dtime), "CanopyFluxes (no allocation)")

    def land_step_dev(self, ds, filterc_dev=None, nc=None, filterp_dev=None, np_=None, dtime=1800.0, structs=None):
        """One land step with canopy on a DeviceState (enqueues only)."""
        if structs is None:
            cs, ps = ds.structs()
            can = ds.canopy_struct()
        else:
            cs, ps, can = structs
        nc = ds.dims.ncol if nc is None else nc
        np_ = ds.dims.npatch if np_ is None else np_
        fc = C.c_void_p(filterc_dev) if filterc_dev is not None else None
        fp = C.c_void_p(filterp_dev) if filterp_dev is not None else None
        self._check(self.L.ctsm_land_step_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ps), C.byref(can),
                                              nc, fc, np_, fp, dtime), "land step")

    # ---- SoilFluxes and the soil energy balance check ---------------------
    def soilfluxes(self, st, filter_c=None, filter_p=None, dtime=1800.0):
        """Drop-in SoilFluxes on a host state (after SoilTemperature)."""
        cs, ps = st.structs()
        sf = st.soilflux_struct()
        fc, fcp, nc = _filter(filter_c, st.dims.ncol)
        fp, fpp, np_ = _filter(filter_p, st.dims.npatch)
        self._check(self.L.ctsm_soilfluxes(self.h, C.byref(st.layout()), C.byref(cs), C.byref(ps), C.byref(sf),
                                           nc, fcp, np_, fpp, dtime), "SoilFluxes")

    def soil_balance_check(self, st, filter_c=None):
        """(max |errsoi_col|, its column) over host arrays."""
        sf = st.soilflux_struct()
        fc, fcp, nc = _filter(filter_c, st.dims.ncol)
        m, idx = C.c_double(), C.c_int32()
        self._check(self.L.ctsm_soil_balance_check(self.h, C.byref(st.layout()), C.byref(sf), nc, fcp, C.byref(m),
                                                   C.byref(idx)), "soil balance check")
        return m.value, idx.value

    def soilfluxes_dev(self, ds, filterc_dev=None, nc=None, filterp_dev=None, np_=None, dtime=1800.0):
        cs, ps = ds.structs()
        sf = ds.soilflux_struct()
        nc = ds.dims.ncol if nc is None else nc
        np_ = ds.dims.npatch if np_ is None else np_
        fc = C.c_void_p(filterc_dev) if filterc_dev is not None else None
        fp = C.c_void_p(filterp_dev) if filterp_dev is not None else None
        self._check(self.L.ctsm_soilfluxes_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ps), C.byref(sf),
                                               nc, fc, np_, fp, dtime), "SoilFluxes")

    def soil_balance_check_dev(self, ds, filterc_dev=None, nc=None):
        sf = ds.soilflux_struct()
        nc = ds.dims.ncol if nc is None else nc
        fc = C.c_void_p(filterc_dev) if filterc_dev is not None else None
        m, idx = C.c_double(), C.c_int32()
        self._check(self.L.ctsm_soil_balance_check_dev(self.h, C.byref(ds.layout()), C.byref(sf), nc, fc,
                                                       C.byref(m), C.byref(idx)), "soil balance check")
        return m.value, idx.value

    def land_step_full_dev(self, ds, filterc_dev=None, nc=None, filterp_dev=None, np_=None, dtime=1800.0,
                           structs=None, hydro=None, water_table=False, drainage=False):
        """Land step with the before-step save, SoilFluxes and, given hydro
        (HydroParams), the surface water input chain; with water_table, the
        water tables and RenewCondensation after SoilWater (enqueues only)."""
        if structs is None:
            cs, ps = ds.structs()
            can, sf = ds.canopy_struct(), ds.soilflux_struct()
        else:
            cs, ps, can, sf = structs
        sw = ds.surfwater_struct() if hydro is not None else None
        wt = ds.watertable_struct() if water_table else None
        dr = ds.drainage_struct() if drainage else None
        nc = ds.dims.ncol if nc is None else nc
        np_ = ds.dims.npatch if np_ is None else np_
        fc = C.c_void_p(filterc_dev) if filterc_dev is not None else None
        fp = C.c_void_p(filterp_dev) if filterp_dev is not None else None
        self._check(self.L.ctsm_land_step_full_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ps),
                                                   C.byref(can), C.byref(sf),
                                                   C.byref(sw) if sw is not None else None,
                                                   C.byref(hydro) if hydro is not None else None,
                                                   C.byref(wt) if wt is not None else None,
                                                   C.byref(dr) if dr is not None else None,
                                                   nc, fc, np_, fp, dtime), "land step (full)")

    def surface_water(self, st, hydro, filter_c=None, dtime=1800.0):
        """Drop-in surface water input chain on a host state."""
        cs, _ = st.structs()
        sw = st.surfwater_struct()
        f, fp, n = _filter(filter_c, st.dims.ncol)
        self._check(self.L.ctsm_surface_water(self.h, C.byref(st.layout()), C.byref(cs), C.byref(sw),
                                              C.byref(hydro), n, fp, dtime), "surface water")

    def surface_water_dev(self, ds, hydro, filter_dev=None, n=None, dtime=1800.0):
        cs, _ = ds.structs()
        sw = ds.surfwater_struct()
        n = ds.dims.ncol if n is None else n
        fp = C.c_void_p(filter_dev) if filter_dev is not None else None
        self._check(self.L.ctsm_surface_water_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(sw),
                                                  C.byref(hydro), n, fp, dtime), "surface water")

    def water_table(self, st, filter_c=None, dtime=1800.0):
        """Drop-in water tables and RenewCondensation on a host state."""
        cs, _ = st.structs()
        wt = st.watertable_struct()
        f, fp, n = _filter(filter_c, st.dims.ncol)
        self._check(self.L.ctsm_water_table(self.h, C.byref(st.layout()), C.byref(cs), C.byref(wt), n, fp, dtime),
                    "water table")

    def water_table_dev(self, ds, filter_dev=None, n=None, dtime=1800.0):
        cs, _ = ds.structs()
        wt = ds.watertable_struct()
        n = ds.dims.ncol if n is None else n
        fp = C.c_void_p(filter_dev) if filter_dev is not None else None
        self._check(self.L.ctsm_water_table_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(wt), n, fp,
                                                dtime), "water table")

    def set_phs_params(self, prm):
        """ctsm_set_phs_params with an abi.PhsParams over host arrays."""
        self._check(self.L.ctsm_set_phs_params(self.h, C.byref(prm)), "set_phs_params")

    def phs_soil_root_conductance(self, lay, phs, n, filt=None, dev=False):
        """Root-soil conductances of PhotosynthesisHydraulicStress over n patches
        (filt: host int32 array, or a device pointer with dev=True; None = the
        first n patches of the layout).  phs holds host arrays, or device arrays
        with dev=True."""
        fn = self.L.ctsm_phs_soil_root_conductance_dev if dev else self.L.ctsm_phs_soil_root_conductance
        fp = C.c_void_p(filt) if dev and filt is not None else (None if filt is None else filt.ctypes.data_as(C.c_void_p))
        self._check(fn(self.h, C.byref(lay), C.byref(phs), n, fp), "phs soil-root conductance")

    def phs_root_sink(self, lay, phs, n, filt=None, dev=False):
        """Compute_EffecRootFrac_And_VertTranSink_HydStress over n columns (as above)."""
        fn = self.L.ctsm_phs_root_sink_dev if dev else self.L.ctsm_phs_root_sink
        fp = C.c_void_p(filt) if dev and filt is not None else (None if filt is None else filt.ctypes.data_as(C.c_void_p))
        self._check(fn(self.h, C.byref(lay), C.byref(phs), n, fp), "phs root sink")

    def phs_calcstress(self, lay, st, n, filt=None, dev=False):
        """calcstress over n patches of a ctsm_phs_stress_state (as above)."""
        fn = self.L.ctsm_phs_calcstress_dev if dev else self.L.ctsm_phs_calcstress
        fp = C.c_void_p(filt) if dev and filt is not None else (None if filt is None else filt.ctypes.data_as(C.c_void_p))
        self._check(fn(self.h, C.byref(lay), C.byref(st), n, fp), "phs calcstress")

    def drainage(self, st, hydro, filter_c=None, dtime=1800.0):
        """Drop-in drainage of HydrologyDrainage on a host state."""
        cs, _ = st.structs()
        dr = st.drainage_struct()
        f, fp, n = _filter(filter_c, st.dims.ncol)
        self._check(self.L.ctsm_drainage(self.h, C.byref(st.layout()), C.byref(cs), C.byref(dr), C.byref(hydro),
                                         n, fp, dtime), "drainage")

    def drainage_dev(self, ds, hydro, filter_dev=None, n=None, dtime=1800.0):
        cs, _ = ds.structs()
        dr = ds.drainage_struct()
        n = ds.dims.ncol if n is None else n
        fp = C.c_void_p(filter_dev) if filter_dev is not None else None
        self._check(self.L.ctsm_drainage_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(dr),
                                             C.byref(hydro), n, fp, dtime), "drainage")

    def hist_update(self, fields):
        """hist_update_hbuf over an array of abi.HistField (enqueues only)."""
        self._check(self.L.ctsm_hist_update_dev(self.h, fields, len(fields)), "history update")

    def hist_normalize(self, fields):
        self._check(self.L.ctsm_hist_normalize_dev(self.h, fields, len(fields)), "history normalize")

    def hist_zero(self, fields):
        self._check(self.L.ctsm_hist_zero_dev(self.h, fields, len(fields)), "history zero")

    def checkpoint_write(self, path, fields):
        """fields: list of (name, pointer, bytes); waits for the stream first."""
        arr = (abi.CkptField * len(fields))(*[abi.CkptField(n.encode(), p, b) for n, p, b in fields])
        self._check(self.L.ctsm_checkpoint_write(self.h, str(path).encode(), arr, len(fields)), "checkpoint write")

    def checkpoint_read(self, path, fields):
        arr = (abi.CkptField * len(fields))(*[abi.CkptField(n.encode(), p, b) for n, p, b in fields])
        self._check(self.L.ctsm_checkpoint_read(self.h, str(path).encode(), arr, len(fields)), "checkpoint read")

    def graph_capture(self, ds, nsteps=1, canopy=True, filterc_dev=None, nc=None, filterp_dev=None, np_=None,
                      dtime=1800.0, structs=None):
        """Capture `nsteps` land steps (canopy=True) or soil steps into one
        replayable graph over the DeviceState; returns the graph handle."""
        if structs is None:
            cs, ps = ds.structs()
            can = ds.canopy_struct() if canopy else None
        else:
            cs, ps, can = structs
            can = can if canopy else None
        nc = ds.dims.ncol if nc is None else nc
        np_ = (ds.dims.npatch if np_ is None else np_) if canopy else 0
        fc = C.c_void_p(filterc_dev) if filterc_dev is not None else None
        fp = C.c_void_p(filterp_dev) if filterp_dev is not None else None
        g = C.c_void_p()
        self._check(self.L.ctsm_graph_capture(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ps),
                                              C.byref(can) if can is not None else None, nc, fc, np_, fp, dtime,
                                              nsteps, C.byref(g)), "ctsm_graph_capture")
        return g.value

    def graph_capture_full(self, ds, nsteps=1, hydro=None, water_table=False, drainage=False, filterc_dev=None,
                           nc=None, filterp_dev=None, np_=None, dtime=1800.0):
        """Capture `nsteps` full land steps (land_step_full_dev's arguments)
        into one replayable graph; returns the graph handle."""
        cs, ps = ds.structs()
        can, sf = ds.canopy_struct(), ds.soilflux_struct()
        sw = ds.surfwater_struct() if hydro is not None else None
        wt = ds.watertable_struct() if water_table else None
        dr = ds.drainage_struct() if drainage else None
        nc = ds.dims.ncol if nc is None else nc
        np_ = ds.dims.npatch if np_ is None else np_
        fc = C.c_void_p(filterc_dev) if filterc_dev is not None else None
        fp = C.c_void_p(filterp_dev) if filterp_dev is not None else None
        ref = lambda x: C.byref(x) if x is not None else None
        g = C.c_void_p()
        self._check(self.L.ctsm_graph_capture_full(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ps),
                                                   C.byref(can), C.byref(sf), ref(sw), ref(hydro), ref(wt), ref(dr),
                                                   nc, fc, np_, fp, dtime, nsteps, C.byref(g)),
                    "ctsm_graph_capture_full")
        return g.value

    def graph_launch(self, g):
        self._check(self.L.ctsm_graph_launch(self.h, C.c_void_p(g)), "ctsm_graph_launch")

    def graph_destroy(self, g):
        self.L.ctsm_graph_destroy(C.c_void_p(g))

    # ---- device memory / events ------------------------------------------
    def alloc(self, nbytes):
        p = C.c_void_p()
        self._check(self.L.ctsm_dev_alloc(self.h, int(nbytes), C.byref(p)), "ctsm_dev_alloc")
        return p.value

    def free(self, ptr):
        self._check(self.L.ctsm_dev_free(self.h, C.c_void_p(ptr)), "ctsm_dev_free")

    def h2d(self, dptr, arr):
        a = np.ascontiguousarray(arr)
        self._check(self.L.ctsm_memcpy(self.h, C.c_void_p(dptr), a.ctypes.data_as(C.c_void_p), a.nbytes, 1),
                    "ctsm_memcpy")

    def d2h(self, arr, dptr):
        assert arr.flags.c_contiguous
        self._check(self.L.ctsm_memcpy(self.h, arr.ctypes.data_as(C.c_void_p), C.c_void_p(dptr), arr.nbytes, 2),
                    "ctsm_memcpy")

    def h2d_2d(self, dptr, dpitch, arr):
        """Copy a 2-D numpy array (rows = levels) into a device region with row pitch `dpitch` bytes."""
        a = np.ascontiguousarray(arr)
        rows = a.shape[0] if a.ndim == 2 else 1
        w = a.nbytes // rows
        self._check(self.L.ctsm_memcpy2d(self.h, C.c_void_p(dptr), int(dpitch), a.ctypes.data_as(C.c_void_p), w, w,
                                         rows, 1), "ctsm_memcpy2d")

    def d2h_2d(self, arr, dptr, spitch):
        assert arr.flags.c_contiguous
        rows = arr.shape[0] if arr.ndim == 2 else 1
        w = arr.nbytes // rows
        self._check(self.L.ctsm_memcpy2d(self.h, arr.ctypes.data_as(C.c_void_p), w, C.c_void_p(dptr), int(spitch), w,
                                         rows, 2), "ctsm_memcpy2d")

    def copy(self, dst, src, nbytes):
        """Kernel copy device -> device on the library stream (enqueues)."""
        self._check(self.L.ctsm_dev_copy(self.h, C.c_void_p(dst), C.c_void_p(src), int(nbytes)), "ctsm_dev_copy")

    def device_array(self, arr):
        """Copy a numpy array to a new device buffer; returns the pointer."""
        a = np.ascontiguousarray(arr)
        p = self.alloc(a.nbytes)
        self.h2d(p, a)
        return p

    def record(self, ev=None):
        e = C.c_void_p(ev)
        self._check(self.L.ctsm_event_record(self.h, C.byref(e)), "ctsm_event_record")
        return e.value

    def elapsed_ms(self, ev0, ev1):
        ms = C.c_float()
        if self.L.ctsm_event_elapsed(C.c_void_p(ev0), C.c_void_p(ev1), C.byref(ms)) != CTSM_OK:
            raise CtsmError(CTSM_E_HIP, 0, "ctsm_event_elapsed")
        return ms.value

    def destroy_event(self, ev):
        self.L.ctsm_event_destroy(C.c_void_p(ev))

    # ---- device-resident state ------------------------------------------
    def step_dev(self, ds, filter_dev=None, n=None, dtime=1800.0, plain=False, structs=None):
        """One soil step on a DeviceState: fused kernel, or the plain
        one-launch-per-routine sequence.  `filter_dev` is a device pointer
        (int) to int32 column indices or None.  Enqueues only."""
        cs, ps = structs if structs is not None else ds.structs()
        lay = ds.layout()
        n = ds.dims.ncol if n is None else n
        fp = C.c_void_p(filter_dev) if filter_dev is not None else None
        if plain:
            self._check(self.L.ctsm_soiltemperature_dev(self.h, C.byref(lay), C.byref(cs), C.byref(ps), n, fp,
                                                        dtime, 1), "SoilTemperature(plain)")
            self._check(self.L.ctsm_soilwater_fractions_dev(self.h, C.byref(lay), C.byref(cs), n, fp),
                        "SetSoilWaterFractions")
            self._check(self.L.ctsm_soilwater_dev(self.h, C.byref(lay), C.byref(cs), n, fp, dtime, 1),
                        "SoilWater(plain)")
        else:
            self._check(self.L.ctsm_soil_step_fused_dev(self.h, C.byref(lay), C.byref(cs), C.byref(ps), n, fp,
                                                        dtime), "soil step")

    def steps_dev(self, ds, nsteps, filter_dev=None, n=None, dtime=1800.0, structs=None):
        """nsteps fused soil steps in one call (enqueues only)."""
        cs, ps = structs if structs is not None else ds.structs()
        n = ds.dims.ncol if n is None else n
        fp = C.c_void_p(filter_dev) if filter_dev is not None else None
        self._check(self.L.ctsm_soil_steps_fused_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ps), n, fp,
                                                     dtime, nsteps), "soil steps")

    def soilwater_dev(self, ds, filter_dev=None, n=None, dtime=1800.0, plain=False):
        cs, _ = ds.structs()
        n = ds.dims.ncol if n is None else n
        fp = C.c_void_p(filter_dev) if filter_dev is not None else None
        self._check(self.L.ctsm_soilwater_dev(self.h, C.byref(ds.layout()), C.byref(cs), n, fp, dtime,
                                              int(plain)), "SoilWater")

    def soilwater_fractions_dev(self, ds, filter_dev=None, n=None):
        cs, _ = ds.structs()
        n = ds.dims.ncol if n is None else n
        fp = C.c_void_p(filter_dev) if filter_dev is not None else None
        self._check(self.L.ctsm_soilwater_fractions_dev(self.h, C.byref(ds.layout()), C.byref(cs), n, fp),
                    "SetSoilWaterFractions")

    def soiltemperature_dev(self, ds, filter_dev=None, n=None, dtime=1800.0, plain=False):
        cs, ps = ds.structs()
        n = ds.dims.ncol if n is None else n
        fp = C.c_void_p(filter_dev) if filter_dev is not None else None
        self._check(self.L.ctsm_soiltemperature_dev(self.h, C.byref(ds.layout()), C.byref(cs), C.byref(ps), n,
                                                    fp, dtime, int(plain)), "SoilTemperature")

    def sync(self):
        self._check(self.L.ctsm_sync(self.h), "sync")


def _filter(filter_, ncol):
    if filter_ is None:
        return None, None, ncol
    f = np.ascontiguousarray(filter_, dtype=np.int32)
    return f, f.ctypes.data_as(abi.i32p), len(f)
