"""Continuous-batching server for perturbed-parameter ensemble members.

Each request is one ensemble member: a set of columns (with their patches)
generated like the full-column workload, with 30 hydrology / energy
parameters scaled (the soil's Clapp-Hornberger and thermal parameters, the
surface's roughness, emissivity and runoff parameters, and 18 PFT parameters
-- stomatal slope and intercept among them -- on PFT rows of its own), to be
run for a number of time steps.  Members share one device-resident pool (a
`DeviceState`); the library's scheduler (`ctsm_serve_*` in the C-ABI) hands
out slots -- column, patch and PFT-row ranges, so admission is by the pool's
HBM budget -- and steps every running member together with full land steps
(`ctsm_land_step_full_dev`: the before-step save, CanopyFluxes,
SoilTemperature, SoilFluxes, the transpiration sink, the surface water chain,
SoilWater, the water tables and RenewCondensation, the drainage) followed by
the soil energy balance check, without waiting for the device between ticks.
This module moves the members' data: it copies an admitted member into its
slot (renumbered into the pool's indices) and starts it, and copies a member
out once the scheduler reports it done.  `checkpoint` writes the pool and the
resident members' bookkeeping; `EnsembleServer.resume` continues from it, bit
for bit, for long served runs.

Columns are independent, so a member's results do not depend on which other
members share its ticks or on its slot position (tests/test_gpu_serve.py
checks a member in the mix against the same member run alone, bit for bit).
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, lib, restart, synthetic
from .state import Dims, DeviceState, HostState

_CAN_KIND = {n: k for n, _, k in abi.CANOPY_FIELDS}
_SF_KIND = {n: k for n, _, k in abi.SOILFLUX_FIELDS}

# Parameters scaled per member by U(0.5, 1.5), with the bounds that keep them
# physical: (group, field, lower, upper) -- group "c" column, "can" column
# array of the canopy state, "hy" column array of the hydrology state, "pft"
# the member's own rows of the PFT table.
PERTURBED = [
    ("c", "watsat", 0.05, 0.9), ("c", "bsw", 1.0, None), ("c", "sucsat", None, None),
    ("c", "hksat", None, None), ("c", "tkmg", None, None), ("c", "tkdry", None, None),
    ("c", "csol", None, None), ("c", "emg", None, 1.0),
    ("can", "z0mg", None, None), ("hy", "wtfact", None, 1.0), ("hy", "topo_slope", None, None),
    ("hy", "h2osfc_thresh", None, None),
    ("pft", "dleaf", None, None), ("pft", "slatop", None, None), ("pft", "leafcn", None, None),
    ("pft", "flnr", None, None), ("pft", "mbbopt", None, None), ("pft", "medlynslope", None, None),
    ("pft", "medlynintercept", None, None), ("pft", "smpso", None, None), ("pft", "smpsc", None, None),
    ("pft", "z0v_Cr", None, None), ("pft", "z0v_Cs", None, None), ("pft", "z0v_c", None, None),
    ("pft", "z0v_cw", None, None), ("pft", "z0v_LAImax", None, None), ("pft", "dbh", None, None),
    ("pft", "fbw", None, 0.95), ("pft", "nstem", None, None), ("pft", "rstem_per_dbh", None, None),
]


def perturb(st, seed, pft=None):
    """Scale the PERTURBED parameters of a member by U(0.5, 1.5) (one factor
    per parameter and member) and restore the relations between them.
    Returns (factors, the member's PFT table)."""
    r = np.random.default_rng(seed)
    f = r.uniform(0.5, 1.5, len(PERTURBED))
    pft = {k: np.array(v) for k, v in (pft or synthetic.pft_table()).items()}
    for (grp, name, lo, hi), s in zip(PERTURBED, f):
        a = {"c": st.col, "can": st.can, "hy": st.hy, "pft": pft}[grp][name]
        a *= s
        if lo is not None or hi is not None:
            np.clip(a, lo, hi, out=a)
    # the soil matric potential at full stomatal closure stays below the one at opening
    pft["smpsc"][:] = np.minimum(pft["smpsc"], pft["smpso"] - 1000.0)
    # the Meier (2022) roughness solves U/u* = a exp(U/u*) by fixed-point
    # iteration, a = c lt / (4 sqrt(Cs + Cr lt / 2)) (CanopyFluxesMod.F90:916-927);
    # a root exists only for a <= 1/e, and a grows with lt up to z0v_LAImax:
    # z0v_c is held to a <= 0.9/e there (beyond it the reference's own
    # iteration runs to inf and the patch to NaN)
    lt = pft["z0v_LAImax"]
    c_max = 0.9 / np.e * 4.0 * np.sqrt(pft["z0v_Cs"] + 0.5 * pft["z0v_Cr"] * lt) / lt
    pft["z0v_c"][:] = np.minimum(pft["z0v_c"], c_max)
    c = st.col
    NS = st.dims.nlevsno
    c["eff_porosity"][...] = np.maximum(0.01, c["watsat"] - np.minimum(
        c["watsat"], c["h2osoi_ice"][NS:] / (c["dz"][NS:] * 917.0)))
    c["h2osoi_liq"][NS:] = np.minimum(c["h2osoi_liq"][NS:], 0.95 * c["watsat"] * c["dz"][NS:] * 1000.0)
    return f, pft


@dataclass
class Request:
    rid: int
    ncol: int
    steps: int
    arrival: float
    seed: int
    state: HostState = None
    pft: dict = None
    ticket: int = -1
    col0: int = -1
    pat0: int = -1
    pft0: int = -1
    left: int = 0
    admitted: float = math.nan
    done: float = math.nan
    result: dict = field(default_factory=dict)


class EnsembleServer:
    """A resident pool of `cap_cols` columns / `cap_patches` patches driven by
    the library's scheduler (ctsm_serve_*)."""

    def __init__(self, ctx, cap_cols, cap_patches, nlevsno=12, nlevsoi=20, nlevgrnd=25, max_steps_per_tick=8,
                 hydro=None, pool=None, dtime=1800.0):
        self.ctx = ctx
        self.dims = Dims(cap_cols, cap_patches, nlevsno, nlevsoi, nlevgrnd, ngrid=cap_cols)
        if pool is None:
            tmpl = HostState(self.dims, fill_nan=False)
            pool = DeviceState(tmpl, ctx)
            del tmpl
        self.pool = pool
        self.hydro = hydro if hydro is not None else synthetic.hydro_params()
        self.structs = (*self.pool.structs(), self.pool.canopy_struct(), self.pool.soilflux_struct())
        self.npft_base = len(synthetic.pft_table()["dleaf"])
        self.max_steps_per_tick = max_steps_per_tick
        self.dtime = dtime
        self.h = ctx.serve_create(self.pool, self.structs, self.hydro, True, True, self.npft_base,
                                  max_steps_per_tick, dtime)
        self.members = {}               # ticket -> Request
        self.finished = []

    def bytes_per_column(self):
        return self.pool.nbytes() / self.dims.ncol

    def _stats(self):
        """The scheduler's counters (kept from the moment of close())."""
        return self.ctx.serve_stats(self.h) if self.h is not None else self._final

    @property
    def worst_errsoi(self):
        return self._stats()[1]

    @property
    def column_steps(self):
        return self._stats()[0]

    @property
    def worst_energy_errors(self):
        """Largest |errseb| and |errlon| (W/m2) over every served step."""
        s = self._stats()
        return s[4], s[6]

    @property
    def active(self):
        return list(self.members.values())

    # ---- slot transfer ----------------------------------------------------
    def _renumber(self, st, req, sign):
        st.col["patchi"] += sign * req.pat0
        st.col["patchf"] += sign * req.pat0
        st.pch["column"] += sign * req.col0
        st.can["gridcell"] += sign * req.col0
        st.can["itype"] += sign * req.pft0

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

    def submit(self, req):
        """Queue a member (its state and PFT rows already built)."""
        req.ticket = self.ctx.serve_submit(self.h, req.state.dims.ncol, req.state.dims.npatch,
                                           len(req.pft["dleaf"]), req.steps)
        req.left = req.steps
        self.members[req.ticket] = req

    def _upload(self, req):
        self.ctx.set_pft_rows(req.pft0, req.pft)
        self._renumber(req.state, req, +1)
        try:
            self._copy(req, True)
        finally:
            self._renumber(req.state, req, -1)

    def pump(self, wait=False, now=None):
        """One turn of the scheduler: retire completed ticks, admit, enqueue a
        tick; then upload and start admitted members and copy out finished
        ones.  Returns the members that finished in this turn."""
        self.ctx.serve_progress(self.h, wait)
        return self._service(now)

    def _service(self, now=None):
        """Upload and start the admitted members, copy out the finished
        ones; returns the latter."""
        done = []
        for t, req in list(self.members.items()):
            state, c0, p0, f0, _ = self.ctx.serve_poll(self.h, t)
            if state == lib.SERVE_ADMITTED:
                req.col0, req.pat0, req.pft0 = c0, p0, f0
                self._upload(req)
                self.ctx.serve_start(self.h, t)
                req.admitted = time.perf_counter() if now is None else now()
            elif state == lib.SERVE_DONE:
                self._copy(req, False)
                self._renumber(req.state, req, -1)     # back to the member's own numbering
                self.ctx.serve_release(self.h, t)
                del self.members[t]
                req.left = 0
                req.done = time.perf_counter() if now is None else now()
                done.append(req)
        self.finished += done
        return done

    def drain(self):
        """Run every submitted member to completion; returns {rid: state}."""
        out = {}
        while self.members:
            for q in self.pump(wait=True):
                out[q.rid] = q.state
        return out

    def checkpoint(self, path):
        """Write the pool, the resident members' slots and remaining steps and
        the queued members to `path`, once the ticks in flight are done and no
        new one is enqueued (members that finish on the way are copied out
        first).  The server carries on afterwards."""
        self.ctx.serve_progress(self.h, quiesce=True)
        self._service()
        extra = {"active": [], "queued": [], "column_steps": self.column_steps, "worst_errsoi": self.worst_errsoi,
                 "max_steps_per_tick": self.max_steps_per_tick, "dtime": self.dtime}
        for t, r in self.members.items():
            state, c0, p0, f0, left = self.ctx.serve_poll(self.h, t)
            if state == lib.SERVE_QUEUED:
                extra["queued"].append(r.rid)
            elif state == lib.SERVE_RUNNING:
                extra["active"].append([r.rid, c0, p0, f0, r.state.dims.ncol, r.state.dims.npatch,
                                        len(r.pft["dleaf"]), left])
            else:
                raise RuntimeError(f"member {r.rid} in state {state} at a quiesced checkpoint")
        restart.checkpoint(self.ctx, self.pool, path, extra=extra)

    @classmethod
    def resume(cls, ctx, path, requests, hydro=None):
        """A server continuing from a checkpoint.  `requests` maps request ids
        to Request objects whose `state` holds arrays of the member's shape and
        whose `pft` is the member's PFT table (they receive the results when
        the member finishes)."""
        pool, extra = restart.restore(ctx, path)
        d = pool.dims
        srv = cls(ctx, d.ncol, d.npatch, d.nlevsno, d.nlevsoi, d.nlevgrnd, extra["max_steps_per_tick"],
                  hydro=hydro, pool=pool, dtime=extra["dtime"])
        for rid, col0, pat0, pft0, ncol, npatch, npft, left in extra["active"]:
            r = requests[rid]
            assert r.state.dims.ncol == ncol and r.state.dims.npatch == npatch, rid
            r.left = left
            r.ticket = ctx.serve_resume_member(srv.h, ncol, npatch, npft, r.left, col0, pat0, pft0)
            r.col0, r.pat0, r.pft0 = col0, pat0, pft0
            ctx.set_pft_rows(pft0, r.pft)
            srv.members[r.ticket] = r
        for rid in extra.get("queued", []):          # queued again, in their order
            srv.submit(requests[rid])
        return srv

    def close(self):
        """Destroy the scheduler and free the pool; the counters stay
        readable."""
        if self.h is None:
            return
        self._final = self.ctx.serve_stats(self.h)
        self.ctx.serve_destroy(self.h)
        self.h = None
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
    """The member's columns (generated like the full-column workload from its
    seed) with its parameters perturbed; its PFT table goes to req.pft."""
    st = synthetic.with_canopy(synthetic.soil_only(req.ncol, req.seed & 0xFFFFFFFF), req.seed & 0xFFFFFFFF, pmax=pmax)
    _, req.pft = perturb(st, req.seed)
    req.state = st
    return st


def serve(ctx, requests, cap_cols, cap_patches, wall_limit=None, max_steps_per_tick=8):
    """Run the request stream in real time: each member is submitted when it
    arrives and the scheduler is pumped without waiting for the device
    (admission, uploads and copies out overlap the steps in flight); returns
    (server, finished requests, wall seconds).  Times on the requests are
    seconds from the start."""
    srv = EnsembleServer(ctx, cap_cols, cap_patches, max_steps_per_tick=max_steps_per_tick)
    pending = sorted(requests, key=lambda q: q.arrival)
    for q in pending:
        if q.state is None:
            build_member(q)
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0
    i = 0
    while True:
        now = clock()
        while i < len(pending) and pending[i].arrival <= now:
            srv.submit(pending[i])
            i += 1
        if srv.members:
            srv.pump(wait=False, now=clock)
            # nothing to admit or retire yet: let the device run
            if srv.members and ctx.serve_stats(srv.h)[3] >= 2:
                srv.pump(wait=True, now=clock)
        elif i < len(pending):
            time.sleep(min(0.01, max(0.0, pending[i].arrival - now)))
        if i >= len(pending) and not srv.members:
            break
        if wall_limit is not None and now > wall_limit:
            break
    return srv, srv.finished, clock()
