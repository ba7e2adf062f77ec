"""Continuous-batching server for perturbed-parameter ensemble members.

Each request is one ensemble member: a set of columns (with their patches)
generated like the full-column workload, with hydrology / energy parameters
scaled, to be run for a number of time steps.  Members share one
device-resident pool (a `DeviceState` whose column / patch / gridcell ranges
are handed out as slots); admission is by the pool's capacity, i.e. the HBM
budget fixed when the pool is created.  Every tick runs land steps over the
columns of all resident members together -- the full step
(`ctsm_land_step_full_dev`: the before-step save, CanopyFluxes,
SoilTemperature, SoilFluxes, the transpiration sink, the surface water chain,
SoilWater, the water tables and RenewCondensation, the drainage) followed by
the soil energy balance check, or with full=False the bare
`ctsm_land_step_dev`; a member leaves the pool, with its final state copied
back, when its steps are done.  `checkpoint` writes the pool and the resident
members' bookkeeping to a file and `EnsembleServer.resume` continues from it,
bit for bit, for long served runs.

Columns are independent, so a member's results do not depend on which other
members share its ticks or on its slot position (tests/test_gpu_serve.py
checks a member in the mix against the same member run alone, bit for bit).
"""
import bisect
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, restart, synthetic
from .state import Dims, DeviceState, HostState

_CAN_KIND = {n: k for n, _, k in abi.CANOPY_FIELDS}
_SF_KIND = {n: k for n, _, k in abi.SOILFLUX_FIELDS}

# Parameters scaled per member by U(0.5, 1.5), with the bounds that keep them
# physical: (group, field, lower, upper) -- group "c" column, "can" canopy.
PERTURBED = [
    ("c", "watsat", 0.05, 0.9), ("c", "bsw", 1.0, None), ("c", "sucsat", None, None),
    ("c", "hksat", None, None), ("c", "tkmg", None, None), ("c", "tkdry", None, None),
    ("c", "csol", None, None), ("c", "qflx_infl", None, None), ("c", "emg", None, 1.0),
    ("c", "forc_lwrad", None, None), ("c", "eflx_bot", None, None),
    ("can", "z0mg", None, None), ("can", "soilbeta", None, 1.0), ("can", "soilresis", None, None),
    ("can", "htop", 0.1, None), ("can", "elai", 1e-3, None), ("can", "esai", None, None),
    ("can", "laisun", None, None), ("can", "laisha", None, None), ("can", "vcmaxcintsun", None, 1.0),
    ("can", "vcmaxcintsha", None, 1.0), ("can", "sabv", None, None), ("can", "parsun_z", None, None),
    ("can", "parsha_z", None, None), ("can", "fwet", None, 1.0), ("can", "liqcan", None, None),
    ("can", "forc_u", None, None), ("can", "forc_v", None, None), ("can", "dayl", None, 86400.0),
    ("can", "rootfr", None, None),
]


def perturb(st, seed):
    """Scale the PERTURBED parameters of a member by U(0.5, 1.5) (one factor
    per parameter and member) and restore the relations between them."""
    r = np.random.default_rng(seed)
    f = r.uniform(0.5, 1.5, len(PERTURBED))
    for (grp, name, lo, hi), s in zip(PERTURBED, f):
        a = st.col[name] if grp == "c" else st.can[name]
        a *= s
        if lo is not None or hi is not None:
            np.clip(a, lo, hi, out=a)
    c, g = st.col, st.can
    g["laisun"][:] = np.minimum(g["laisun"], g["elai"])
    g["laisha"][:] = g["elai"] - g["laisun"]
    g["laisun_z"][:] = g["laisun"]
    g["laisha_z"][:] = g["laisha"]
    rf = g["rootfr"]
    rf /= np.maximum(rf.sum(axis=0, keepdims=True), 1e-300)
    g["max_dayl"][:] = np.maximum(g["max_dayl"], g["dayl"])
    g["thv"][:] = g["forc_th"] * (1.0 + 0.61 * g["forc_q"])
    NS = st.dims.nlevsno
    c["eff_porosity"][...] = np.maximum(0.01, c["watsat"] - np.minimum(
        c["watsat"], c["h2osoi_ice"][NS:] / (c["dz"][NS:] * 917.0)))
    c["h2osoi_liq"][NS:] = np.minimum(c["h2osoi_liq"][NS:], 0.95 * c["watsat"] * c["dz"][NS:] * 1000.0)
    return f


@dataclass
class Request:
    rid: int
    ncol: int
    steps: int
    arrival: float
    seed: int
    state: HostState = None
    col0: int = -1
    pat0: int = -1
    left: int = 0
    admitted: float = math.nan
    done: float = math.nan
    result: dict = field(default_factory=dict)


class RangeAlloc:
    """First-fit allocator of contiguous index ranges [0, cap)."""

    def __init__(self, cap):
        self.cap = cap
        self.free = [(0, cap)]          # sorted (start, length)

    def alloc(self, n):
        for i, (s, l) in enumerate(self.free):
            if l >= n:
                if l == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + n, l - n)
                return s
        return -1

    def take(self, s, n):
        """Mark [s, s + n) as used (it must lie inside one free range)."""
        for i, (a, l) in enumerate(self.free):
            if a <= s and s + n <= a + l:
                self.free[i:i + 1] = [r for r in ((a, s - a), (s + n, a + l - s - n)) if r[1] > 0]
                return
        raise ValueError(f"range [{s}, {s + n}) is not free")

    def release(self, s, n):
        i = bisect.bisect(self.free, (s, 0))
        self.free.insert(i, (s, n))
        merged = []
        for a, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == a:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((a, l))
        self.free = merged


class EnsembleServer:
    """A resident pool of `cap_cols` columns / `cap_patches` patches."""

    def __init__(self, ctx, cap_cols, cap_patches, nlevsno=12, nlevsoi=20, nlevgrnd=25, max_steps_per_tick=8,
                 full=True, hydro=None, pool=None):
        self.ctx = ctx
        self.dims = Dims(cap_cols, cap_patches, nlevsno, nlevsoi, nlevgrnd, ngrid=cap_cols)
        if pool is None:
            tmpl = HostState(self.dims, fill_nan=False)
            pool = DeviceState(tmpl, ctx)
            del tmpl
        self.pool = pool
        self.full = full
        self.hydro = hydro if hydro is not None else synthetic.hydro_params()
        self.structs = (*self.pool.structs(), self.pool.canopy_struct())
        if full:
            self.structs = (*self.structs, self.pool.soilflux_struct())
        self.worst_errsoi = 0.0
        self.cols = RangeAlloc(cap_cols)
        self.pats = RangeAlloc(cap_patches)
        self.active = []
        self.fc = ctx.alloc(4 * cap_cols)
        self.fp = ctx.alloc(4 * cap_patches)
        self.nc = self.np = 0
        self.dirty = True
        self.max_steps_per_tick = max_steps_per_tick
        self.column_steps = 0

    def bytes_per_column(self):
        return self.pool.nbytes() / self.dims.ncol

    # ---- slot transfer ----------------------------------------------------
    def _copy(self, req, to_pool):
        st, d = req.state, self.dims
        for group, host, dev in (("c", st.col, self.pool.col), ("p", st.pch, self.pool.pch),
                                 ("g", st.can, self.pool.can), ("s", st.sf, self.pool.sf),
                                 ("y", st.hy, self.pool.hy)):
            for k, a in host.items():
                shape, dt = self.pool._shapes[(group, k)]
                isp = (group == "p" or (group == "g" and _CAN_KIND[k] in ("p", "soip"))
                       or (group == "s" and _SF_KIND[k] == "p"))
                off = req.pat0 if isp else req.col0
                cap = d.npatch if isp else d.ncol
                esz = np.dtype(dt).itemsize
                ptr = dev[k] + off * esz
                if to_pool:
                    self.ctx.h2d_2d(ptr, cap * esz, a)
                else:
                    self.ctx.d2h_2d(a, ptr, cap * esz)

    def admit(self, req):
        st = req.state
        c0 = self.cols.alloc(st.dims.ncol)
        if c0 < 0:
            return False
        p0 = self.pats.alloc(st.dims.npatch)
        if p0 < 0:
            self.cols.release(c0, st.dims.ncol)
            return False
        req.col0, req.pat0 = c0, p0
        # renumber into the pool's 1-based global indices
        st.col["patchi"] += p0
        st.col["patchf"] += p0
        st.pch["column"] += c0
        st.can["gridcell"] += c0
        self._copy(req, True)
        st.col["patchi"] -= p0
        st.col["patchf"] -= p0
        st.pch["column"] -= c0
        st.can["gridcell"] -= c0
        req.left = req.steps
        self.active.append(req)
        self.dirty = True
        return True

    def retire(self, req):
        self._copy(req, False)
        st = req.state                      # back to the member's own 1-based numbering
        st.col["patchi"] -= req.pat0
        st.col["patchf"] -= req.pat0
        st.pch["column"] -= req.col0
        st.can["gridcell"] -= req.col0
        self.cols.release(req.col0, req.state.dims.ncol)
        self.pats.release(req.pat0, req.state.dims.npatch)
        self.active.remove(req)
        self.dirty = True

    def _filters(self):
        if not self.dirty:
            return
        fc = np.concatenate([np.arange(r.col0 + 1, r.col0 + r.state.dims.ncol + 1, dtype=np.int32)
                             for r in self.active]) if self.active else np.zeros(0, np.int32)
        fp = np.concatenate([np.arange(r.pat0 + 1, r.pat0 + r.state.dims.npatch + 1, dtype=np.int32)
                             for r in self.active]) if self.active else np.zeros(0, np.int32)
        if len(fc):
            self.ctx.h2d(self.fc, fc)
            self.ctx.h2d(self.fp, fp)
        self.nc, self.np = len(fc), len(fp)
        self.dirty = False

    def tick(self, dtime=1800.0):
        """Run k land steps over every resident member (k = the fewest steps
        any member still needs, at most max_steps_per_tick); returns k."""
        if not self.active:
            return 0
        self._filters()
        k = min(self.max_steps_per_tick, min(r.left for r in self.active))
        for _ in range(k):
            if self.full:
                self.ctx.land_step_full_dev(self.pool, filterc_dev=self.fc, nc=self.nc, filterp_dev=self.fp,
                                            np_=self.np, dtime=dtime, structs=self.structs, hydro=self.hydro,
                                            water_table=True, drainage=True)
                m, _ = self.ctx.soil_balance_check_dev(self.pool, filterc_dev=self.fc, nc=self.nc)
                self.worst_errsoi = max(self.worst_errsoi, m)
            else:
                self.ctx.land_step_dev(self.pool, filterc_dev=self.fc, nc=self.nc, filterp_dev=self.fp, np_=self.np,
                                       dtime=dtime, structs=self.structs)
        self.ctx.sync()
        for r in self.active:
            r.left -= k
        self.column_steps += k * self.nc
        return k

    def checkpoint(self, path):
        """Write the pool (after the enqueued steps) and the resident members'
        slots and remaining steps to `path`."""
        extra = {"active": [[r.rid, r.col0, r.pat0, r.left, r.state.dims.ncol, r.state.dims.npatch]
                            for r in self.active],
                 "column_steps": self.column_steps, "worst_errsoi": self.worst_errsoi,
                 "max_steps_per_tick": self.max_steps_per_tick, "full": self.full}
        restart.checkpoint(self.ctx, self.pool, path, extra=extra)

    @classmethod
    def resume(cls, ctx, path, requests, hydro=None):
        """A server continuing from a checkpoint.  `requests` maps request ids
        to Request objects whose `state` holds arrays of the member's shape
        (they receive the results when the member retires)."""
        pool, extra = restart.restore(ctx, path)
        d = pool.dims
        srv = cls(ctx, d.ncol, d.npatch, d.nlevsno, d.nlevsoi, d.nlevgrnd, extra["max_steps_per_tick"],
                  full=extra["full"], hydro=hydro, pool=pool)
        for rid, col0, pat0, left, ncol, npatch in extra["active"]:
            r = requests[rid]
            assert r.state.dims.ncol == ncol and r.state.dims.npatch == npatch, rid
            srv.cols.take(col0, ncol)
            srv.pats.take(pat0, npatch)
            r.col0, r.pat0, r.left = col0, pat0, left
            srv.active.append(r)
        srv.column_steps = extra["column_steps"]
        srv.worst_errsoi = extra["worst_errsoi"]
        return srv

    def close(self):
        self.ctx.free(self.fc)
        self.ctx.free(self.fp)
        self.pool.free()


def make_requests(n, rate, seed, max_cols=20000, max_days=365, steps_per_day=48, day_scale=1.0):
    """Poisson arrivals at `rate` per second; sizes log-uniform in
    [1, max_cols] columns; run lengths log-uniform in [1, max_days] days of
    steps_per_day steps, multiplied by day_scale (1 = the full length)."""
    r = np.random.default_rng(seed)
    t = np.cumsum(r.exponential(1.0 / rate, n))
    size = np.exp(r.uniform(0.0, math.log(max_cols), n)).astype(int).clip(1, max_cols)
    days = np.exp(r.uniform(0.0, math.log(max_days), n))
    steps = np.maximum(1, np.round(days * steps_per_day * day_scale)).astype(int)
    return [Request(i, int(size[i]), int(steps[i]), float(t[i]), int(seed * 1000003 + i)) for i in range(n)]


def build_member(req, pmax=16):
    st = synthetic.with_canopy(synthetic.soil_only(req.ncol, req.seed & 0xFFFFFFFF), req.seed & 0xFFFFFFFF, pmax=pmax)
    perturb(st, req.seed)
    req.state = st
    return st


def serve(ctx, requests, cap_cols, cap_patches, wall_limit=None, max_steps_per_tick=8, full=True):
    """Run the request stream in real time; returns (server, finished requests, wall seconds)."""
    srv = EnsembleServer(ctx, cap_cols, cap_patches, max_steps_per_tick=max_steps_per_tick, full=full)
    pending = sorted(requests, key=lambda q: q.arrival)
    queue, finished = [], []
    for q in pending:
        build_member(q)
    t0 = time.perf_counter()
    i = 0
    while True:
        now = time.perf_counter() - t0
        while i < len(pending) and pending[i].arrival <= now:
            queue.append(pending[i])
            i += 1
        still = []
        for q in queue:                     # FIFO admission while it fits
            if not still and srv.admit(q):
                q.admitted = time.perf_counter() - t0
            else:
                still.append(q)
        queue = still
        if srv.active:
            srv.tick()
            for q in [q for q in srv.active if q.left <= 0]:
                srv.retire(q)
                q.done = time.perf_counter() - t0
                finished.append(q)
        elif i < len(pending):
            time.sleep(min(0.01, max(0.0, pending[i].arrival - now)))
        if i >= len(pending) and not queue and not srv.active:
            break
        if wall_limit is not None and now > wall_limit:
            break
    wall = time.perf_counter() - t0
    return srv, finished, wall
