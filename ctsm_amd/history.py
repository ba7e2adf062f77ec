"""History tapes of device-resident state.

The reference accumulates every history field into a buffer each step
(hist_update_hbuf, histFileMod.F90:1362-2169), divides the averages before a
write (hfields_normalize, :2302-2358) and clears the buffers after it
(hfields_zero, :2361-2381).  Here the buffers live in HBM next to the state:
each step costs one kernel launch over all fields (ctsm_hist_update_dev), and
only a history write copies anything to the host.  Fields are kept at the
resolution of their source (column or patch, with levels), as the reference
does when it does not average up to the gridcell; files are netCDF-3
(scipy.io.netcdf_file) with the reference's field names, units and spval.
"""

import numpy as np

from . import abi
from .state import CANOPY_ALIAS, SOILFLUX_ALIAS, SURFWATER_ALIAS

SPVAL = 1.e36

# name, group, state field, levels, avgflag, units, long name.  Levels: None
# for a 1-D field, ("sno", first, count) or ("soi", first, count) with
# count "levgrnd" / "levsoi" for the two soil level sets (1-based first level).
DEFAULT_FIELDS = [
    ("TSOI", "c", "t_soisno", ("sno", 1, "levgrnd"), "A", "K", "soil temperature"),
    ("SOILLIQ", "c", "h2osoi_liq", ("sno", 1, "levsoi"), "A", "kg/m2", "soil liquid water"),
    ("SOILICE", "c", "h2osoi_ice", ("sno", 1, "levsoi"), "A", "kg/m2", "soil ice"),
    ("H2OSFC", "c", "h2osfc", None, "A", "mm", "surface water depth"),
    ("TG", "c", "t_grnd", None, "A", "K", "ground temperature"),
    ("QINFL", "c", "qflx_infl", None, "A", "mm/s", "infiltration"),
    ("QOVER", "y", "qflx_surf", None, "A", "mm/s", "total surface runoff"),
    ("FSAT", "y", "fsat", None, "A", "unitless", "fractional area with water table at surface"),
    ("ZWT", "y", "zwt", None, "A", "m", "water table depth"),
    ("ZWT_PERCH", "y", "zwt_perched", None, "A", "m", "perched water table depth"),
    ("FROST_TABLE", "y", "frost_table", None, "A", "m", "frost table depth"),
    ("ERRSOI", "s", "errsoi_col", None, "A", "W/m^2", "soil/lake energy conservation error"),
    ("FSH", "s", "eflx_sh_tot", None, "A", "W/m^2", "sensible heat"),
    ("EFLX_LH_TOT", "s", "eflx_lh_tot", None, "A", "W/m^2", "total latent heat flux [+ to atm]"),
    ("TV", "g", "t_veg", None, "A", "K", "vegetation temperature"),
]


def _kind_tables():
    t = {}
    for grp, fields in (("c", abi.COL_FIELDS), ("p", abi.PATCH_FIELDS), ("g", abi.CANOPY_FIELDS),
                        ("s", abi.SOILFLUX_FIELDS), ("y", abi.SURFWATER_FIELDS + abi.WATERTABLE_FIELDS)):
        for n, typ, kind in fields:
            t[(grp, n)] = (typ, kind)
    return t


_KINDS = _kind_tables()


def device_field(ds, group, key):
    """Device pointer and (type, kind) of a state field, aliases resolved."""
    typ, kind = _KINDS[(group, key)]
    g, k = group, key
    for _ in range(3):
        if g == "g" and k in CANOPY_ALIAS:
            g, k = CANOPY_ALIAS[k]
        elif g == "s" and k in SOILFLUX_ALIAS:
            g, k = SOILFLUX_ALIAS[k]
        elif g == "y" and k in SURFWATER_ALIAS:
            g, k = SURFWATER_ALIAS[k]
    store = {"c": ds.col, "p": ds.pch, "g": ds.can, "s": ds.sf, "y": ds.hy}[g]
    return store[k], typ, kind


class HistoryTape:
    """One history tape over a DeviceState: buffers in HBM, updated each step."""

    def __init__(self, ctx, ds, fields=DEFAULT_FIELDS):
        self.ctx, self.ds = ctx, ds
        d = ds.dims
        self.specs, self.bufs = [], []
        desc = []
        for name, group, key, levels, flag, units, long_name in fields:
            ptr, typ, kind = device_field(ds, group, key)
            if typ != "f64":
                raise ValueError(f"{name}: history fields are FP64")
            patch = kind in ("p", "soip", "lyr")
            n1d = d.npatch if patch else d.ncol
            if levels is None:
                first_off, num2d = 0, 1
            else:
                lk, first, count = levels
                num2d = {"levgrnd": d.nlevgrnd, "levsoi": d.nlevsoi}.get(count, count)
                first_off = (d.nlevsno + first - 1) if lk == "sno" else (first - 1)
            active = device_field(ds, "g", "active")[0] if patch and ds.can else None
            hbuf = ctx.alloc(max(8 * n1d * num2d, 8))
            nacs = ctx.alloc(max(4 * n1d * num2d, 4))
            self.bufs += [hbuf, nacs]
            desc.append(abi.HistField(ptr + 8 * first_off * n1d, active, hbuf, nacs, n1d, n1d, num2d, ord(flag), 0))
            self.specs.append((name, "pft" if patch else "column", levels, num2d, flag, units, long_name))
        self.desc = (abi.HistField * len(desc))(*desc)
        self.nsteps = 0
        ctx.hist_zero(self.desc)

    def update(self):
        """hist_update_hbuf for every field (enqueues only)."""
        self.ctx.hist_update(self.desc)
        self.nsteps += 1

    def buffers(self):
        """Host copies of (hbuf, nacs) per field, shaped (num2d, n1d) or (n1d,)."""
        self.ctx.sync()
        out = {}
        for (name, _, levels, num2d, _, _, _), f in zip(self.specs, self.desc):
            h = np.empty((num2d, f.n1d))
            n = np.empty((num2d, f.n1d), dtype=np.int32)
            self.ctx.d2h(h, f.hbuf)
            self.ctx.d2h(n, f.nacs)
            out[name] = (h if levels else h[0], n if levels else n[0])
        return out

    def write(self, path, time=0.0):
        """Normalise, write a netCDF-3 file, clear the buffers (one D2H per field)."""
        from scipy.io import netcdf_file
        self.ctx.hist_normalize(self.desc)
        data = self.buffers()
        d = self.ds.dims
        with netcdf_file(str(path), "w", version=2) as f:
            f.createDimension("time", None)
            f.createDimension("column", d.ncol)
            f.createDimension("pft", d.npatch)
            f.createDimension("levgrnd", d.nlevgrnd)
            f.createDimension("levsoi", d.nlevsoi)
            t = f.createVariable("time", "d", ("time",))
            t[0] = time
            f.nsteps_averaged = self.nsteps
            for name, dim1, levels, num2d, flag, units, long_name in self.specs:
                dims = ("time", dim1) if levels is None else ("time", levels[2], dim1)
                v = f.createVariable(name, "d", dims)
                v[0] = data[name][0]
                v.units = units
                v.long_name = long_name
                v.cell_methods = {"A": "time: mean", "S": "time: sum", "I": "time: point",
                                  "X": "time: maximum", "M": "time: minimum"}[flag]
                v.missing_value = SPVAL
                v._FillValue = SPVAL
        self.ctx.hist_zero(self.desc)
        self.nsteps = 0

    def free(self):
        for b in self.bufs:
            self.ctx.free(b)
        self.bufs = []
