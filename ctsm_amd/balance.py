"""Inputs of the column water balance check and the water mass
(ctsm_water_balance_state, include/ctsm_gpu.h): the column stores beyond the
soil / canopy state, the step's water fluxes and begwb / endwb.

The reference's BalanceCheck (BalanceCheckMod.F90:588-630) reads these from
waterstate / waterflux / waterbalance types that routines outside this
library's path fill (CanopyHydrology, SnowHydrology, the atmosphere
forcing).  WaterBalance holds them as numpy arrays (host struct for the
drop-in style calls); DeviceWaterBalance keeps device copies and the struct
of device pointers that ctsm_water_mass_dev / ctsm_water_balance_check_dev
take.
"""
import ctypes as C

import numpy as np

from . import abi

FLUXES = ("forc_rain", "forc_snow", "qflx_floodc", "qflx_sfc_irrig", "qflx_glcice_dyn_water_flux", "qflx_evap_tot",
          "qflx_surf", "qflx_qrgwl", "qflx_drain", "qflx_drain_perched", "qflx_ice_runoff",
          "qflx_snwcp_discarded_liq", "qflx_snwcp_discarded_ice")
# the fluxes that add water to the column (BalanceCheckMod.F90:593-597); the rest remove it
GAINS = ("forc_rain", "forc_snow", "qflx_floodc", "qflx_sfc_irrig", "qflx_glcice_dyn_water_flux")

AQUIFER_WATER_BASELINE = 5000.0        # clm_varcon.F90:124 (mm)


class WaterBalance:
    """Host arrays of a ctsm_water_balance_state; a member left out (or
    None) is passed as NULL."""

    def __init__(self, arrays):
        self.a = {}
        types = {n: t for n, t, _ in abi.WATERBAL_FIELDS}
        for name, arr in arrays.items():
            if arr is None:
                continue
            if name not in types:
                raise KeyError(name)
            dt = np.int32 if types[name] == "i32" else np.float64
            self.a[name] = np.ascontiguousarray(arr, dtype=dt)

    def __getitem__(self, k):
        return self.a[k]

    def __setitem__(self, k, v):
        self.a[k] = v

    def copy(self):
        return WaterBalance({k: v.copy() for k, v in self.a.items()})

    def struct(self):
        w = abi.WaterBalanceState()
        for name, t, _ in abi.WATERBAL_FIELDS:
            if name in self.a:
                setattr(w, name, self.a[name].ctypes.data_as(abi._CT[t]))
        return w


class DeviceWaterBalance:
    """Device copies of a WaterBalance."""

    def __init__(self, ctx, wb):
        self.ctx, self.wb, self.ptr = ctx, wb, {}
        for name, arr in wb.a.items():
            p = ctx.alloc(max(arr.nbytes, 8))
            ctx.h2d(p, arr)
            self.ptr[name] = p

    def struct(self):
        w = abi.WaterBalanceState()
        for name, t, _ in abi.WATERBAL_FIELDS:
            if name in self.ptr:
                setattr(w, name, C.cast(C.c_void_p(self.ptr[name]), abi._CT[t]))
        return w

    def fetch(self, name):
        out = np.empty_like(self.wb.a[name])
        self.ctx.sync()
        self.ctx.d2h(out, self.ptr[name])
        return out

    def free(self):
        for p in self.ptr.values():
            self.ctx.free(p)
        self.ptr = {}


def synthetic(st, seed, excess_ice=True):
    """Seeded stores and fluxes for the columns of a state `st` (ctsm_amd.state
    host state): an aquifer near its baseline, a tenth of the columns
    hydrologically inactive and a twentieth inactive, fluxes of a wet day
    (mm/s), begwb / endwb of the size of a soil column's water.  Also sets
    the state's unlayered snow (st.col["h2osno_no_layers"]) on columns
    without snow layers."""
    d = st.dims
    n = d.ncol
    rng = np.random.default_rng(seed)
    snl = st.col["snl"]
    st.col["h2osno_no_layers"][:] = np.where(snl == 0, rng.uniform(0.0, 8.0, n) * (rng.random(n) < 0.4), 0.0)
    a = {
        "total_plant_stored_h2o": rng.uniform(0.0, 0.5, n),
        "wa": AQUIFER_WATER_BASELINE + rng.normal(0.0, 200.0, n),
        "excess_ice": rng.uniform(0.0, 30.0, (d.nlevgrnd, n)) * (rng.random((d.nlevgrnd, n)) < 0.2)
        if excess_ice else None,
        "dynbal_baseline_liq": rng.normal(0.0, 5.0, n),
        "dynbal_baseline_ice": rng.normal(0.0, 5.0, n),
        "hydrologically_active": (rng.random(n) >= 0.1).astype(np.int32),
        "col_active": (rng.random(n) >= 0.05).astype(np.int32),
        "begwb": rng.uniform(500.0, 9000.0, n),
        "errh2o": np.zeros(n),
    }
    for f in FLUXES:
        a[f] = rng.uniform(0.0, 2e-4, n) * (rng.random(n) < 0.7)
    a["endwb"] = a["begwb"] + rng.normal(0.0, 0.5, n)
    return WaterBalance(a)
