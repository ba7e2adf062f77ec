"""Host side of the plant-hydraulic-stress (PHS) soil entries.

PhsArrays holds the members of ctsm_phs_state as numpy arrays laid out as the
header says (patch or column fastest, levels slowest); host_struct() points a
ctsm_phs_state at them for the drop-in calls, DevicePhs uploads them and points
one at HBM for the _dev calls.  phs_params() builds ctsm_phs_params from the
PFT tables (psi50 / ck given as (4, npft): segments sun, sha, xyl, root).
synthetic() draws a case over the ranges CTSM produces (20SL_8.5m layers,
roots thinning with depth, soils from wet to past the vulnerability cut).
"""
import ctypes as C

import numpy as np

from . import abi

_NLEV = {"c": lambda d: 1, "p": lambda d: 1, "soi": lambda d: d["nlevgrnd"], "soip": lambda d: d["nlevgrnd"],
         "sno": lambda d: d["nlevsno"] + d["nlevgrnd"], "veg": lambda d: 4}
_PATCH_KINDS = ("p", "soip", "veg")


class PhsArrays:
    """Numpy arrays of every ctsm_phs_state member plus the layout numbers."""

    FIELDS, STRUCT = abi.PHS_FIELDS, abi.PhsState

    def __init__(self, ncol, npatch, nlevsno=12, nlevsoi=20, nlevgrnd=25, begc=1, begp=1):
        self.ncol, self.npatch, self.begc, self.begp = ncol, npatch, begc, begp
        self.dims = {"nlevsno": nlevsno, "nlevsoi": nlevsoi, "nlevgrnd": nlevgrnd}
        self.a = {}
        for name, typ, kind in self.FIELDS:
            n = npatch if kind in _PATCH_KINDS else ncol
            shape = (n,) if kind in ("c", "p") else (_NLEV[kind](self.dims), n)
            self.a[name] = np.zeros(shape, dtype=np.int32 if typ == "i32" else np.float64)

    def __getitem__(self, k):
        return self.a[k]

    def layout(self):
        return abi.Layout(begc_alloc=self.begc, ldc=self.ncol, begp_alloc=self.begp, ldp=self.npatch,
                          begg_alloc=1, ldg=1)

    def copy(self):
        o = type(self).__new__(type(self))
        o.ncol, o.npatch, o.begc, o.begp, o.dims = self.ncol, self.npatch, self.begc, self.begp, dict(self.dims)
        o.a = {k: v.copy() for k, v in self.a.items()}
        return o

    def host_struct(self, omit=()):
        s = self.STRUCT()
        for name, typ, _ in self.FIELDS:
            if name in omit:
                continue
            setattr(s, name, self.a[name].ctypes.data_as(abi.i32p if typ == "i32" else abi.f64p))
        return s


class DevicePhs:
    """Device copies of a PhsArrays (one allocation per member)."""

    def __init__(self, ctx, arrs):
        self.ctx, self.arrs, self.ptr = ctx, arrs, {}
        for name, arr in arrs.a.items():
            p = ctx.alloc(max(arr.nbytes, 8))
            ctx.h2d(p, arr)
            self.ptr[name] = p

    def struct(self, omit=()):
        s = self.arrs.STRUCT()
        for name, typ, _ in self.arrs.FIELDS:
            if name not in omit:
                setattr(s, name, C.cast(C.c_void_p(self.ptr[name]), abi.i32p if typ == "i32" else abi.f64p))
        return s

    def fetch(self):
        """Host copies of every member after the device calls."""
        out = self.arrs.copy()
        self.ctx.sync()
        for name, arr in out.a.items():
            self.ctx.d2h(arr, self.ptr[name])
        return out

    def free(self):
        for p in self.ptr.values():
            self.ctx.free(p)
        self.ptr = {}


class StressArrays(PhsArrays):
    """Numpy arrays of every ctsm_phs_stress_state member (calcstress)."""
    FIELDS, STRUCT = abi.PHS_STRESS_FIELDS, abi.PhsStressState


def phs_params(krmax, root_radius, root_density, froot_leaf, psi50, ck, kmax=None, crop=None):
    """ctsm_phs_params over contiguous copies (kept alive on the returned object);
    psi50, ck and kmax as (4, npft); crop (npft) 0 / 1 or None."""
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (krmax, root_radius, root_density, froot_leaf)]
    segs = [np.ascontiguousarray(x, dtype=np.float64).reshape(4, -1) for x in (psi50, ck)]
    if kmax is not None:
        segs.append(np.ascontiguousarray(kmax, dtype=np.float64).reshape(4, -1))
    npft = arrs[0].size
    if any(a.size != npft for a in arrs) or any(a.shape[1] != npft for a in segs):
        raise ValueError("PHS tables must all have npft entries")
    prm = abi.PhsParams(npft=npft, pad=0)
    for name, a in zip(("krmax", "root_radius", "root_density", "froot_leaf", "psi50", "ck", "kmax"), arrs + segs):
        setattr(prm, name, a.ctypes.data_as(abi.f64p))
    keep = arrs + segs
    if crop is not None:
        c = np.ascontiguousarray(crop, dtype=np.float64)
        if c.size != npft:
            raise ValueError("crop must have npft entries")
        prm.crop = c.ctypes.data_as(abi.f64p)
        keep.append(c)
    prm._keep = keep
    return prm


def default_params(npft=17, crop=None):
    """PHS tables of the magnitudes in CTSM's parameter file: krmax 2e-9 .. 1e-8,
    root radius 0.29 mm, root density 0.31e6 g/m3, psi50 from -150000 to
    -530000 mm, ck 3.95; crop (npft) marks crop types (None: no crop table)."""
    rng = np.random.default_rng(11)
    krmax = rng.uniform(2e-9, 1e-8, npft)
    psi50 = -rng.uniform(1.5e5, 5.3e5, (4, npft))
    ck = np.full((4, npft), 3.95)
    kmax = rng.uniform(1e-8, 4e-8, (4, npft))
    return phs_params(krmax, np.full(npft, 0.29e-3), np.full(npft, 0.31e6), rng.uniform(1.0, 2.0, npft), psi50, ck,
                      kmax, crop)


def synthetic(ncol, max_patches=4, seed=0, npft=17, nlevsno=12, nlevsoi=20, nlevgrnd=25):
    """A case of ncol columns with 0..max_patches patches each."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(0, max_patches + 1, ncol)
    counts[0] = max(counts[0], 1)
    npatch = int(counts.sum())
    A = PhsArrays(ncol, npatch, nlevsno, nlevsoi, nlevgrnd)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    A["patchi"][:] = A.begp + first
    A["patchf"][:] = A.begp + first + counts - 1
    A["column"][:] = np.repeat(np.arange(A.begc, A.begc + ncol), counts)
    # 20SL_8.5m-like soil: thickness growing with depth, nodes at the layer middles
    dz = np.concatenate([np.full(4, 0.02), np.full(4, 0.04), np.full(6, 0.06), np.full(6, 0.12),
                         np.full(nlevgrnd - 20, 1.0)])[:nlevgrnd]
    zi = np.concatenate([[0.0], np.cumsum(dz)])
    z = 0.5 * (zi[:-1] + zi[1:])
    A["dz"][nlevsno:, :] = dz[:, None] * rng.uniform(0.9, 1.1, (1, ncol))
    A["z"][nlevsno:, :] = z[:, None] * rng.uniform(0.9, 1.1, (1, ncol))
    A["hksat"][:] = rng.uniform(1e-3, 2e-2, (nlevgrnd, ncol))
    A["hk_l"][:] = A["hksat"] * 10.0 ** rng.uniform(-8, 0.3, (nlevgrnd, ncol))
    A["smp_l"][:] = -10.0 ** rng.uniform(1.5, 6.2, (nlevgrnd, ncol))
    A["itype"][:] = rng.integers(0, npft, npatch)
    A["active"][:] = rng.random(npatch) < 0.9
    A["frac_veg_nosno"][:] = rng.random(npatch) < 0.85
    w = rng.random(npatch) * (rng.random(npatch) < 0.95)
    A["wtcol"][:] = w
    A["froot_carbon"][:] = rng.uniform(0, 600, npatch) * (rng.random(npatch) < 0.9)
    A["tsai"][:] = rng.uniform(0, 2, npatch)
    A["tlai"][:] = rng.uniform(0, 6, npatch)
    depth = rng.integers(5, nlevsoi + 1, npatch)
    prof = np.exp(-rng.uniform(1, 6, npatch)[None, :] * z[:, None])
    prof[np.arange(nlevgrnd)[:, None] >= depth[None, :]] = 0.0
    A["rootfr"][:] = prof / prof.sum(0, keepdims=True)
    A["vegwp"][:] = -rng.uniform(1e3, 3e5, (4, npatch))
    A["k_soil_root"][:] = rng.uniform(0, 1e-9, (nlevgrnd, npatch))
    return A


def synthetic_stress(ncol, max_patches=4, seed=0, npft=17, nlevsno=12, nlevsoi=20, nlevgrnd=25):
    """A calcstress case: day and night patches (a positive vegwp(sun) marks
    night), sunlit and shaded leaf area from none to dense, unstressed
    conductances of a ci solve, soils from wet to dry, root conductances of the
    magnitudes the soil-side kernel gives."""
    rng = np.random.default_rng(seed)
    counts = rng.integers(1, max_patches + 1, ncol)
    npatch = int(counts.sum())
    A = StressArrays(ncol, npatch, nlevsno, nlevsoi, nlevgrnd)
    A["column"][:] = np.repeat(np.arange(A.begc, A.begc + ncol), counts)
    A["itype"][:] = rng.integers(0, npft, npatch)
    A["local_morning"][:] = rng.random(npatch) < 0.5
    lai = rng.uniform(0, 6, npatch) * (rng.random(npatch) < 0.9)
    fsun = rng.uniform(0, 0.6, npatch)
    A["laisun"][:] = lai * fsun * (rng.random(npatch) < 0.9)
    A["laisha"][:] = lai * (1 - fsun) * (rng.random(npatch) < 0.95)
    A["elai"][:] = A["laisun"] + A["laisha"]
    A["esai"][:] = rng.uniform(0, 1.5, npatch)
    A["tsai"][:] = A["esai"] * rng.uniform(1.0, 1.2, npatch)
    A["htop"][:] = rng.uniform(0.5, 30, npatch)
    A["fdry"][:] = rng.uniform(0.3, 1.0, npatch)
    A["thm"][:] = rng.uniform(270, 310, npatch)
    A["gb_mol"][:] = rng.uniform(2e5, 3e6, npatch)
    A["gs_mol_sun"][:] = rng.uniform(0, 4e5, npatch) * (rng.random(npatch) < 0.9)
    A["gs_mol_sha"][:] = rng.uniform(0, 2e5, npatch) * (rng.random(npatch) < 0.9)
    A["qsatl"][:] = rng.uniform(0.005, 0.03, npatch)
    A["qaf"][:] = A["qsatl"] - rng.uniform(-0.002, 0.012, npatch)
    ns = nlevsno
    dz = np.concatenate([np.full(4, 0.02), np.full(4, 0.04), np.full(6, 0.06), np.full(6, 0.12),
                         np.full(nlevgrnd - 20, 1.0)])[:nlevgrnd]
    zi = np.concatenate([[0.0], np.cumsum(dz)])
    A["z"][ns:, :] = (0.5 * (zi[:-1] + zi[1:]))[:, None]
    A["smp_l"][:] = -10.0 ** rng.uniform(2, 5.3, (nlevgrnd, ncol))
    A["forc_rho"][:] = rng.uniform(1.1, 1.3, ncol)
    A["forc_pbot"][:] = rng.uniform(8.5e4, 1.02e5, ncol)
    k = 10.0 ** rng.uniform(-11, -8.5, (nlevgrnd, npatch))
    k[0] = 0.0
    k[nlevsoi:] = 0.0
    k[:, rng.random(npatch) < 0.03] = 0.0                     # no roots at all
    A["k_soil_root"][:] = k
    base = -rng.uniform(2e4, 2e5, npatch)
    A["vegwp"][:] = base[None, :] - np.array([3e4, 2.5e4, 1e4, 0.0])[:, None] * rng.uniform(0, 1, (1, npatch))
    night = rng.random(npatch) < 0.3
    A["vegwp"][0, night] = -A["vegwp"][0, night]              # the reference's night mark
    A["bsun"][:] = -1.0
    A["bsha"][:] = -1.0
    A["qflx_tran_veg"][:] = -1.0
    return A


# ---- CanopyFluxes with plant hydraulic stress ---------------------------------

# ctsm_phs_state members that are arrays of the land state (group, name)
_PHS_ALIAS = {"itype": ("can", "itype"), "column": ("pch", "column"), "active": ("can", "active"),
              "frac_veg_nosno": ("pch", "frac_veg_nosno"), "wtcol": ("pch", "wtcol"), "rootfr": ("can", "rootfr"),
              "patchi": ("col", "patchi"), "patchf": ("col", "patchf"), "dz": ("col", "dz"), "z": ("col", "z"),
              "hksat": ("col", "hksat"), "hk_l": ("col", "hk_l"), "smp_l": ("col", "smp_l"),
              "qflx_rootsoi": ("col", "qflx_rootsoi")}
_NLEV2 = dict(_NLEV, phs2=lambda d: 2)


class CanopyPhs:
    """The arrays CanopyFluxes with plant hydraulic stress needs beyond a land
    state (HostState): the conductance members of ctsm_phs_state the state
    does not hold (froot_carbon, tsai, tlai, vegwp, the three conductances,
    the sink outputs) and every ctsm_phs_canopy_state member.  vegwp is one
    array for both structs."""

    def __init__(self, dims):
        self.dims = {"nlevsno": dims.nlevsno, "nlevsoi": dims.nlevsoi, "nlevgrnd": dims.nlevgrnd}
        self.npatch, self.ncol = dims.npatch, dims.ncol
        self.a = {}
        for name, typ, kind in abi.PHS_FIELDS + abi.PHS_CANOPY_FIELDS:
            if name in _PHS_ALIAS or name in self.a:
                continue
            n = self.npatch if kind in _PATCH_KINDS + ("phs2",) else self.ncol
            shape = (n,) if kind in ("c", "p") else (_NLEV2[kind](self.dims), n)
            self.a[name] = np.zeros(shape, dtype=np.int32 if typ == "i32" else np.float64)

    def __getitem__(self, k):
        return self.a[k]

    def copy(self):
        o = CanopyPhs.__new__(CanopyPhs)
        o.dims, o.npatch, o.ncol = dict(self.dims), self.npatch, self.ncol
        o.a = {k: v.copy() for k, v in self.a.items()}
        return o

    def _ptr(self, st, name, typ):
        if name in _PHS_ALIAS:
            g, k = _PHS_ALIAS[name]
            arr = getattr(st, g)[k]
        else:
            arr = self.a[name]
        assert arr.flags.c_contiguous
        return arr.ctypes.data_as(abi.i32p if typ == "i32" else abi.f64p)

    def host_structs(self, st):
        """(ctsm_phs_state, ctsm_phs_canopy_state) over st's and these arrays."""
        phs, phc = abi.PhsState(), abi.PhsCanopyState()
        for name, typ, _ in abi.PHS_FIELDS:
            setattr(phs, name, self._ptr(st, name, typ))
        for name, typ, _ in abi.PHS_CANOPY_FIELDS:
            setattr(phc, name, self._ptr(st, name, typ))
        return phs, phc


def canopy_phs(st, seed=0, night_frac=0.3):
    """PHS inputs for a land state of the full-column workload: fine-root
    carbon, stem and leaf area without snow burial, the layer's leaf area,
    local morning flags, vegetation water potentials from a cold start
    (-2.5e4 mm) to stressed, previous stress factors; the soil's matric
    potential and conductivity set from its liquid water by Clapp and
    Hornberger (the values SoilWater leaves).  night_frac of the patches get
    no PAR (the night branch).  Modifies st's smp_l, hk_l and PAR; returns a
    CanopyPhs."""
    rng = np.random.default_rng(seed)
    d = st.dims
    A = CanopyPhs(d)
    np_, ns = d.npatch, d.nlevsno
    can, col = st.can, st.col
    A["froot_carbon"][:] = rng.uniform(0, 600, np_) * (rng.random(np_) < 0.9)
    A["tsai"][:] = can["esai"] * rng.uniform(1.0, 1.2, np_)
    A["tlai"][:] = can["elai"] * rng.uniform(1.0, 1.2, np_)
    A["tlai_z"][:] = can["elai"]
    A["local_morning"][:] = rng.random(np_) < 0.5
    base = -rng.uniform(2.0e4, 1.5e5, np_)
    A["vegwp"][:] = base[None, :] - np.array([2e4, 1.5e4, 5e3, 0.0])[:, None] * rng.uniform(0, 1, (1, np_))
    A["bsun"][:] = rng.uniform(0.2, 1.0, np_)
    A["bsha"][:] = rng.uniform(0.2, 1.0, np_)
    for k in ("vegwp_ln", "vegwp_pd", "an_sun", "an_sha", "vcmax_z_phs", "tpu_z_phs", "kp_z_phs", "ac_phs",
              "aj_phs", "ap_phs", "ag_phs"):
        A[k][:] = np.nan
    ng = d.nlevgrnd
    liq = col["h2osoi_liq"][ns:ns + ng]
    vol = liq / (col["dz"][ns:ns + ng] * 1000.0)
    s = np.clip(vol / col["watsat"], 0.01, 1.0)
    col["smp_l"][:] = -col["sucsat"] * s ** (-col["bsw"])
    col["hk_l"][:] = col["hksat"] * s ** (2.0 * col["bsw"] + 3.0)
    dark = rng.random(np_) < night_frac
    can["parsun_z"][dark] = 0.0
    can["parsha_z"][dark] = 0.0
    return A


class DeviceCanopyPhs:
    """Device copies of a CanopyPhs, with structs that point at a
    DeviceState's arrays for the aliased members."""

    def __init__(self, ctx, arrs, ds):
        self.ctx, self.arrs, self.ds, self.ptr = ctx, arrs, ds, {}
        for name, arr in arrs.a.items():
            p = ctx.alloc(max(arr.nbytes, 8))
            ctx.h2d(p, arr)
            self.ptr[name] = p

    def _dptr(self, name, typ):
        if name in _PHS_ALIAS:
            g, k = _PHS_ALIAS[name]
            if g == "can":
                from .state import CANOPY_ALIAS
                if k in CANOPY_ALIAS:
                    g2, k2 = CANOPY_ALIAS[k]
                    p = self.ds.pch[k2] if g2 == "p" else self.ds.col[k2]
                else:
                    p = self.ds.can[k]
            else:
                p = getattr(self.ds, g)[k]
        else:
            p = self.ptr[name]
        return C.cast(C.c_void_p(p), abi.i32p if typ == "i32" else abi.f64p)

    def structs(self):
        phs, phc = abi.PhsState(), abi.PhsCanopyState()
        for name, typ, _ in abi.PHS_FIELDS:
            setattr(phs, name, self._dptr(name, typ))
        for name, typ, _ in abi.PHS_CANOPY_FIELDS:
            setattr(phc, name, self._dptr(name, typ))
        return phs, phc

    def fetch(self):
        out = self.arrs.copy()
        self.ctx.sync()
        for name, arr in out.a.items():
            self.ctx.d2h(arr, self.ptr[name])
        return out

    def free(self):
        for p in self.ptr.values():
            self.ctx.free(p)
        self.ptr = {}
