"""Column / patch state containers in the reference's memory layout.

A state is a dict of arrays, one per field of ctsm_col_state /
ctsm_patch_state, laid out exactly as the reference allocates them
(ColumnType.F90:137-155 and friends): level-major, column-fastest, so array
[k, c] is Fortran element (c, lo + k).  Host states hold numpy arrays; device
states hold buffers allocated through the library's own C-ABI.
"""
import ctypes as C

import numpy as np

from . import abi

_NP = {"i32": np.int32, "f64": np.float64}


# Canopy fields that are the same arrays as a column or patch field.
CANOPY_ALIAS = {
    "column": ("p", "column"), "wtcol": ("p", "wtcol"), "frac_veg_nosno": ("p", "frac_veg_nosno"),
    "dlrad": ("p", "dlrad"), "cgrnd": ("p", "cgrnd"), "eflx_sh_grnd": ("p", "eflx_sh_grnd"),
    "eflx_sh_snow": ("p", "eflx_sh_snow"), "eflx_sh_soil": ("p", "eflx_sh_soil"),
    "eflx_sh_h2osfc": ("p", "eflx_sh_h2osfc"), "qflx_evap_soi": ("p", "qflx_evap_soi"),
    "qflx_ev_snow": ("p", "qflx_ev_snow"), "qflx_ev_soil": ("p", "qflx_ev_soil"),
    "qflx_ev_h2osfc": ("p", "qflx_ev_h2osfc"), "qflx_rootsoi_col": ("c", "qflx_rootsoi"),
}


# SoilFluxes fields that are the same arrays as a column, patch or canopy field
# (the rest -- the saved temperatures, errsoi and the flux outputs -- are its own).
SOILFLUX_ALIAS = {
    "active": ("g", "active"), "eflx_sh_grnd": ("p", "eflx_sh_grnd"), "qflx_evap_soi": ("p", "qflx_evap_soi"),
    "qflx_ev_snow": ("p", "qflx_ev_snow"), "qflx_ev_soil": ("p", "qflx_ev_soil"),
    "qflx_ev_h2osfc": ("p", "qflx_ev_h2osfc"), "cgrnds": ("g", "cgrnds"), "cgrndl": ("g", "cgrndl"),
    "ulrad": ("g", "ulrad"), "eflx_sh_veg": ("g", "eflx_sh_veg"), "eflx_sh_stem": ("g", "eflx_sh_stem"),
    "qflx_evap_veg": ("g", "qflx_evap_veg"), "qflx_tran_veg": ("g", "qflx_tran_veg"), "t_skin": ("g", "t_skin"),
}
SOILFLUX_OWN = [(n, t, k) for n, t, k in abi.SOILFLUX_FIELDS if n not in SOILFLUX_ALIAS]

# surface-water fields that are column-state arrays; the rest are its own
SURFWATER_ALIAS = {"h2osfc": ("c", "h2osfc"), "qflx_infl": ("c", "qflx_infl")}
SURFWATER_OWN = [(n, t, k) for n, t, k in abi.SURFWATER_FIELDS if n not in SURFWATER_ALIAS]
# water-table fields: zwt, zwt_perched and frost_table are the surface water's
# (kept in the same group); h2osoi_vol and the column dew fluxes are its own
WATERTABLE_SHARED = ("zwt", "zwt_perched", "frost_table")
WATERTABLE_OWN = [(n, t, k) for n, t, k in abi.WATERTABLE_FIELDS if n not in WATERTABLE_SHARED]
# drainage fields: the surface water's and water table's are shared (same group)
DRAINAGE_OWN = [(n, t, k) for n, t, k in abi.DRAINAGE_FIELDS
                if n not in {f[0] for f in abi.SURFWATER_FIELDS + abi.WATERTABLE_FIELDS}]


class Dims:
    def __init__(self, ncol, npatch, nlevsno=12, nlevsoi=20, nlevgrnd=25, ngrid=0):
        self.ncol, self.npatch, self.ngrid = int(ncol), int(npatch), int(ngrid)
        self.nlevsno, self.nlevsoi, self.nlevgrnd = nlevsno, nlevsoi, nlevgrnd

    def shape(self, kind):
        n = {"p": self.npatch, "lyr": self.npatch, "soip": self.npatch, "g": self.ngrid}.get(kind, self.ncol)
        nl = abi.nlev_of(kind, self.nlevsno, self.nlevgrnd)
        return (n,) if kind in ("c", "p", "g") else (nl, n)


class HostState:
    """numpy arrays for every column and patch field."""

    def __init__(self, dims, fill_nan=True):
        self.dims = dims
        self.col = {}
        self.pch = {}
        for name, t, kind in abi.COL_FIELDS:
            a = np.zeros(dims.shape(kind), dtype=_NP[t])
            if fill_nan and t == "f64":
                a[...] = np.nan
            self.col[name] = a
        for name, t, kind in abi.PATCH_FIELDS:
            a = np.zeros(dims.shape(kind), dtype=_NP[t])
            if fill_nan and t == "f64":
                a[...] = np.nan
            self.pch[name] = a
        self.can = {}
        self.sf = {}
        if dims.ngrid > 0:
            for name, t, kind in abi.CANOPY_FIELDS:
                if name in CANOPY_ALIAS:
                    continue
                a = np.zeros(dims.shape(kind), dtype=_NP[t])
                if fill_nan and t == "f64":
                    a[...] = np.nan
                self.can[name] = a
            for name, t, kind in SOILFLUX_OWN:
                a = np.zeros(dims.shape(kind), dtype=_NP[t])
                if fill_nan and t == "f64":
                    a[...] = np.nan
                self.sf[name] = a
        self.hy = {}
        for name, t, kind in SURFWATER_OWN + WATERTABLE_OWN + DRAINAGE_OWN:
            a = np.zeros(dims.shape(kind), dtype=_NP[t])
            if fill_nan and t == "f64":
                a[...] = np.nan
            self.hy[name] = a

    def copy(self):
        o = HostState.__new__(HostState)
        o.dims = self.dims
        o.col = {k: v.copy() for k, v in self.col.items()}
        o.pch = {k: v.copy() for k, v in self.pch.items()}
        o.can = {k: v.copy() for k, v in self.can.items()}
        o.sf = {k: v.copy() for k, v in getattr(self, "sf", {}).items()}
        o.hy = {k: v.copy() for k, v in getattr(self, "hy", {}).items()}
        return o

    def canopy_array(self, name):
        """The array behind canopy field `name` (aliases resolve to col/pch)."""
        if name in CANOPY_ALIAS:
            g, k = CANOPY_ALIAS[name]
            return self.pch[k] if g == "p" else self.col[k]
        return self.can[name]

    def layout(self):
        return abi.Layout(1, self.dims.ncol, 1, self.dims.npatch, 1, self.dims.ngrid)

    def canopy_struct(self):
        cs = abi.CanopyState()
        for name, t, _ in abi.CANOPY_FIELDS:
            a = self.canopy_array(name)
            assert a.flags.c_contiguous
            setattr(cs, name, a.ctypes.data_as(abi._CT[t]))
        return cs

    def soilflux_array(self, name):
        """The array behind SoilFluxes field `name` (aliases resolve to col/pch/can)."""
        if name in SOILFLUX_ALIAS:
            g, k = SOILFLUX_ALIAS[name]
            if g == "p":
                return self.pch[k]
            if g == "c":
                return self.col[k]
            return self.canopy_array(k)
        return self.sf[name]

    def soilflux_struct(self):
        ss = abi.SoilFluxState()
        for name, t, _ in abi.SOILFLUX_FIELDS:
            a = self.soilflux_array(name)
            assert a.flags.c_contiguous
            setattr(ss, name, a.ctypes.data_as(abi._CT[t]))
        return ss

    def surfwater_struct(self):
        sw = abi.SurfaceWaterState()
        for name, t, _ in abi.SURFWATER_FIELDS:
            a = self.col[SURFWATER_ALIAS[name][1]] if name in SURFWATER_ALIAS else self.hy[name]
            assert a.flags.c_contiguous
            setattr(sw, name, a.ctypes.data_as(abi._CT[t]))
        return sw

    def watertable_struct(self):
        wt = abi.WaterTableState()
        for name, t, _ in abi.WATERTABLE_FIELDS:
            a = self.hy[name]
            assert a.flags.c_contiguous
            setattr(wt, name, a.ctypes.data_as(abi._CT[t]))
        return wt

    def drainage_struct(self):
        dr = abi.DrainageState()
        for name, t, _ in abi.DRAINAGE_FIELDS:
            a = self.col[SURFWATER_ALIAS[name][1]] if name in SURFWATER_ALIAS else self.hy[name]
            assert a.flags.c_contiguous
            setattr(dr, name, a.ctypes.data_as(abi._CT[t]))
        return dr

    def structs(self):
        """ctypes ColState / PatchState pointing at these numpy buffers."""
        cs = abi.ColState()
        for name, t, _ in abi.COL_FIELDS:
            a = self.col[name]
            assert a.flags.c_contiguous
            setattr(cs, name, a.ctypes.data_as(abi._CT[t]))
        ps = abi.PatchState()
        for name, t, _ in abi.PATCH_FIELDS:
            a = self.pch[name]
            assert a.flags.c_contiguous
            setattr(ps, name, a.ctypes.data_as(abi._CT[t]))
        return cs, ps

    def subset(self, cols):
        """State holding only columns `cols` (0-based, sorted) and their patches,
        renumbered densely (used by the alone-vs-batched determinism tests)."""
        cols = np.asarray(cols)
        d = self.dims
        pi = self.col["patchi"][cols] - 1
        pf = self.col["patchf"][cols] - 1
        pidx = np.concatenate([np.arange(a, b + 1) for a, b in zip(pi, pf)])
        nd = Dims(len(cols), len(pidx), d.nlevsno, d.nlevsoi, d.nlevgrnd)
        o = HostState(nd)
        for name, t, kind in abi.COL_FIELDS:
            src = self.col[name]
            o.col[name] = np.ascontiguousarray(src[..., cols])
        for name, t, kind in abi.PATCH_FIELDS:
            src = self.pch[name]
            o.pch[name] = np.ascontiguousarray(src[..., pidx])
        npc = pf - pi + 1
        start = np.concatenate([[0], np.cumsum(npc)[:-1]])
        o.col["patchi"] = (start + 1).astype(np.int32)
        o.col["patchf"] = (start + npc).astype(np.int32)
        o.pch["column"] = np.repeat(np.arange(1, len(cols) + 1, dtype=np.int32), npc)
        o.can = {}
        if d.ngrid > 0:
            # synthetic states carry one gridcell per column
            o.dims.ngrid = len(cols)
            for name, t, kind in abi.CANOPY_FIELDS:
                if name in CANOPY_ALIAS:
                    continue
                src = self.can[name]
                idx = {"p": pidx, "soip": pidx, "g": cols}.get(kind, cols)
                o.can[name] = np.ascontiguousarray(src[..., idx])
            o.can["gridcell"] = o.pch["column"].copy()
            o.sf = {}
            for name, t, kind in SOILFLUX_OWN:
                idx = pidx if kind == "p" else cols
                o.sf[name] = np.ascontiguousarray(self.sf[name][..., idx])
        o.hy = {k: np.ascontiguousarray(v[..., cols]) for k, v in getattr(self, "hy", {}).items()}
        return o


class DeviceState:
    """Device-resident mirror of a HostState, allocated through the library
    (ctsm_dev_alloc / ctsm_memcpy), so no other GPU runtime is involved."""

    def __init__(self, host, ctx):
        self.ctx = ctx
        self.dims = host.dims
        self._shapes = {}
        self.col, self.pch, self.can, self.sf, self.hy = {}, {}, {}, {}, {}
        for group, src, dst in (("c", host.col, self.col), ("p", host.pch, self.pch),
                                ("g", getattr(host, "can", {}), self.can), ("s", getattr(host, "sf", {}), self.sf),
                                ("y", getattr(host, "hy", {}), self.hy)):
            for k, v in src.items():
                a = np.ascontiguousarray(v)
                ptr = ctx.alloc(max(a.nbytes, 8))
                ctx.h2d(ptr, a)
                dst[k] = ptr
                self._shapes[(group, k)] = (a.shape, a.dtype)

    def layout(self):
        return abi.Layout(1, self.dims.ncol, 1, self.dims.npatch, 1, self.dims.ngrid)

    def canopy_struct(self):
        cs = abi.CanopyState()
        for name, t, _ in abi.CANOPY_FIELDS:
            if name in CANOPY_ALIAS:
                g, k = CANOPY_ALIAS[name]
                ptr = self.pch[k] if g == "p" else self.col[k]
            else:
                ptr = self.can[name]
            setattr(cs, name, C.cast(C.c_void_p(ptr), abi._CT[t]))
        return cs

    def soilflux_struct(self):
        ss = abi.SoilFluxState()
        for name, t, _ in abi.SOILFLUX_FIELDS:
            if name in SOILFLUX_ALIAS:
                g, k = SOILFLUX_ALIAS[name]
                if g == "g" and k in CANOPY_ALIAS:
                    g, k = CANOPY_ALIAS[k]
                ptr = {"p": self.pch, "c": self.col, "g": self.can}[g][k]
            else:
                ptr = self.sf[name]
            setattr(ss, name, C.cast(C.c_void_p(ptr), abi._CT[t]))
        return ss

    def surfwater_struct(self):
        sw = abi.SurfaceWaterState()
        for name, t, _ in abi.SURFWATER_FIELDS:
            ptr = self.col[SURFWATER_ALIAS[name][1]] if name in SURFWATER_ALIAS else self.hy[name]
            setattr(sw, name, C.cast(C.c_void_p(ptr), abi._CT[t]))
        return sw

    def watertable_struct(self):
        wt = abi.WaterTableState()
        for name, t, _ in abi.WATERTABLE_FIELDS:
            setattr(wt, name, C.cast(C.c_void_p(self.hy[name]), abi._CT[t]))
        return wt

    def drainage_struct(self):
        dr = abi.DrainageState()
        for name, t, _ in abi.DRAINAGE_FIELDS:
            ptr = self.col[SURFWATER_ALIAS[name][1]] if name in SURFWATER_ALIAS else self.hy[name]
            setattr(dr, name, C.cast(C.c_void_p(ptr), abi._CT[t]))
        return dr

    def structs(self):
        cs = abi.ColState()
        for name, t, _ in abi.COL_FIELDS:
            setattr(cs, name, C.cast(C.c_void_p(self.col[name]), abi._CT[t]))
        ps = abi.PatchState()
        for name, t, _ in abi.PATCH_FIELDS:
            setattr(ps, name, C.cast(C.c_void_p(self.pch[name]), abi._CT[t]))
        return cs, ps

    def to_host(self):
        o = HostState.__new__(HostState)
        o.dims = self.dims
        o.col, o.pch, o.can, o.sf, o.hy = {}, {}, {}, {}, {}
        for group, src, dst in (("c", self.col, o.col), ("p", self.pch, o.pch), ("g", self.can, o.can),
                                ("s", self.sf, o.sf), ("y", self.hy, o.hy)):
            for k, ptr in src.items():
                shape, dt = self._shapes[(group, k)]
                a = np.empty(shape, dtype=dt)
                self.ctx.d2h(a, ptr)
                dst[k] = a
        return o

    def upload(self, host):
        """Overwrite the device copy with a host state of the same dims."""
        for group, src, dst in (("c", host.col, self.col), ("p", host.pch, self.pch),
                                ("g", getattr(host, "can", {}), self.can), ("s", getattr(host, "sf", {}), self.sf),
                                ("y", getattr(host, "hy", {}), self.hy)):
            for k, v in src.items():
                self.ctx.h2d(dst[k], np.ascontiguousarray(v))

    def nbytes(self):
        return sum(int(np.prod(s)) * np.dtype(d).itemsize for s, d in self._shapes.values())

    def free(self):
        for d in (self.col, self.pch, self.can, self.sf, self.hy):
            for k, ptr in d.items():
                self.ctx.free(ptr)
            d.clear()
