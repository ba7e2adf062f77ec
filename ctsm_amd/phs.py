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

    def __init__(self, ncol, npatch, nlevsno=12, nlevsoi=20, nlevgrnd=25, begc=1, begp=1):
        self.ncol, self.npatch, self.begc, self.begp = ncol, npatch, begc, begp
        self.dims = {"nlevsno": nlevsno, "nlevsoi": nlevsoi, "nlevgrnd": nlevgrnd}
        self.a = {}
        for name, typ, kind in abi.PHS_FIELDS:
            n = npatch if kind in _PATCH_KINDS else ncol
            shape = (n,) if kind in ("c", "p") else (_NLEV[kind](self.dims), n)
            self.a[name] = np.zeros(shape, dtype=np.int32 if typ == "i32" else np.float64)

    def __getitem__(self, k):
        return self.a[k]

    def layout(self):
        return abi.Layout(begc_alloc=self.begc, ldc=self.ncol, begp_alloc=self.begp, ldp=self.npatch,
                          begg_alloc=1, ldg=1)

    def copy(self):
        o = PhsArrays.__new__(PhsArrays)
        o.ncol, o.npatch, o.begc, o.begp, o.dims = self.ncol, self.npatch, self.begc, self.begp, dict(self.dims)
        o.a = {k: v.copy() for k, v in self.a.items()}
        return o

    def host_struct(self, omit=()):
        s = abi.PhsState()
        for name, typ, _ in abi.PHS_FIELDS:
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
        s = abi.PhsState()
        for name, typ, _ in abi.PHS_FIELDS:
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


def phs_params(krmax, root_radius, root_density, froot_leaf, psi50, ck):
    """ctsm_phs_params over contiguous copies (kept alive on the returned object)."""
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (krmax, root_radius, root_density, froot_leaf)]
    p50 = np.ascontiguousarray(psi50, dtype=np.float64).reshape(4, -1)
    ckk = np.ascontiguousarray(ck, dtype=np.float64).reshape(4, -1)
    npft = arrs[0].size
    if any(a.size != npft for a in arrs) or p50.shape[1] != npft or ckk.shape[1] != npft:
        raise ValueError("PHS tables must all have npft entries")
    prm = abi.PhsParams(npft=npft, pad=0)
    for name, a in zip(("krmax", "root_radius", "root_density", "froot_leaf", "psi50", "ck"), arrs + [p50, ckk]):
        setattr(prm, name, a.ctypes.data_as(abi.f64p))
    prm._keep = arrs + [p50, ckk]
    return prm


def default_params(npft=17):
    """PHS tables of the magnitudes in CTSM's parameter file: krmax 2e-9 .. 1e-8,
    root radius 0.29 mm, root density 0.31e6 g/m3, psi50 from -150000 to
    -530000 mm, ck 3.95."""
    rng = np.random.default_rng(11)
    krmax = rng.uniform(2e-9, 1e-8, npft)
    psi50 = -rng.uniform(1.5e5, 5.3e5, (4, npft))
    ck = np.full((4, npft), 3.95)
    return phs_params(krmax, np.full(npft, 0.29e-3), np.full(npft, 0.31e6), rng.uniform(1.0, 2.0, npft), psi50, ck)


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
