"""Benchmark of the column biogeophysics step on one MI355X.

A "step" is one full land time step (dtime = 1800 s) over every column of the
batch, in clm_drv order: the before-step save, CanopyFluxes (per-patch
leaf-temperature iteration with Photosynthesis, FrictionVelocity, QSat),
SoilTemperature (snow + soil heat solve with phase change) with
SetSoilWaterFractions, SoilFluxes, the transpiration sink, the surface water
input chain, the moisture-form Richards solve (SoilWater), the water tables
with RenewCondensation, and the soil-column drainage.  The headline workload
is 1,048,576 columns (snow layers uniform 0..12, bedrock index uniform
5..20, 1-16 patches per column, FP64, seeded synthetic inputs) resident in
HBM, generated in binned column order, for 48 steps (one day); the soil
launch shape and the CanopyFluxes round split are chosen by the autotuner;
the 48 timed steps are captured once as one HIP graph and launched once.
The same 48 steps are also run in the plain form (one launch per reference
routine, CanopyFluxes iterating every patch to convergence in one launch).

Prints ONE JSON line with the metric, the roofline of the dominant kernel
(algorithmic bytes per CanopyFluxes call / its measured time vs 8 TB/s) and
of the whole step, a CPU baseline (the C oracle's full land step on a
seeded sample of the same workload, one column chunk per thread) and, in
"detail", the soil-only uneven batch (65,536 columns) and the served
ensemble.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Measured HBM traffic (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE, KiB per
# dispatch, scaled by the same counters' readings on a streaming copy of
# known size) and FP64 VALU share (SQ_INSTS_VALU_*_F64 over the launch time
# against the 78.6 TFLOP/s vector peak) of a dominant kernel, keyed by
# (kernel, workload, columns), with the launch shape they were measured at.
# The bench reports them only for a run of that workload, size and launch
# shape; any other run reports null.
PROFILED = {
    # profiles/canopy_rounds_canopy_uneven1048576.md: FETCH 8,222,014 KiB x 2.0 + WRITE 17,863,759 KiB
    # x 1.0 per call; 470.3 GFLOP FP64 in 43.50 ms
    ("k_canopy_round", "canopy_uneven", 1048576): {
        "traffic": (8222014 * 2.0 + 17863759 * 1.0) * 1024,
        "fp64_valu_frac": 470.3e9 / 43.50e-3 / 78.6e12,
        "tuning": {"canopy_first_iterations": 6, "canopy_round_iterations": 6, "canopy_waves": 1},
    },
    # profiles/richards_heat_solve_canopy_uneven1048576.md: the three isolated calls after the steps,
    # FETCH 3,121,768 KiB x 2.0 + WRITE 2,415,621 KiB x 1.0 (Richards), 4,832,447 x 2.0 + 4,272,971
    # (heat solve) per call
    ("k_soilwater_g", "canopy_uneven", 1048576): {
        "traffic": (3121768 * 2.0 + 2415621 * 1.0) * 1024,
        "tuning": {"richards_lanes": 8},          # measured in 128-lane blocks; the block size moves no bytes
    },
    ("k_soiltemp", "canopy_uneven", 1048576): {
        "traffic": (4832447 * 2.0 + 4272971 * 1.0) * 1024,
        "tuning": {},
    },
}


def richards_bytes(st):
    """Compulsory HBM bytes of one Richards launch (k_soilwater*), summed over
    the columns: per Richards layer (nbedrock of them) 10 FP64 read
    (h2osoi_liq, dz, z, watsat, bsw, sucsat, hksat, icefrac, eff_porosity,
    qflx_rootsoi) and 5 written (h2osoi_liq, smp_l, hk_l, qin, qout); per column
    nbedrock + qflx_infl read and qcharge + num_substeps written:
    B_W = 120 L + 28 bytes."""
    L = st.col["nbedrock"].astype(np.int64)
    return int((120 * L + 28).sum())


def soiltemp_bytes(st):
    """Compulsory HBM bytes of one heat-solve launch (SoilTemperature), summed
    over columns:
      per column   : 14 FP64 + 4 int32 scalars read, 18 FP64 scalars written;
      snow layer   : 6 FP64 read (t, liq, ice, dz, z, zi), 7 FP64 + 1 int32
                     written (t, liq, ice, thk, fact, bw + imelt);
      all nlevsno snow slots: qflx_snomelt_lyr, qflx_snofrz_lyr written (2 FP64);
      ground level : 12 FP64 read (t, liq, ice, dz, z, zi, watsat, tkmg, tkdry,
                     csol, bsw, sucsat) + zi(0), 6 FP64 + 1 int32 written
                     (t, liq, ice, thk, fact, eflx_fgr + imelt);
      patch        : 15 FP64 + 1 int32 read, sabg_lyr over snl+1..1 read,
                     3 FP64 written.
    """
    S = -st.col["snl"].astype(np.int64)
    P = (st.col["patchf"] - st.col["patchi"] + 1).astype(np.int64)
    NS, NG = st.dims.nlevsno, st.dims.nlevgrnd
    per_col = (14 * 8 + 4 * 4 + 18 * 8
               + S * (6 * 8 + 7 * 8 + 4)
               + NS * 2 * 8
               + NG * (12 * 8 + 6 * 8 + 4) + 8
               + P * (15 * 8 + 4 + 3 * 8) + P * (S + 1) * 8)
    return int(per_col.sum())


_CANOPY_INOUT = {"displa", "z0mv", "t_veg", "t_stem", "snocan", "liqcan", "cgrnds", "cgrndl"}


def canopy_bytes(st):
    """Compulsory HBM bytes of one CanopyFluxes launch (k_canopy): every patch
    field of ctsm_canopy_state counted once if read and once if written
    (inout both), FP64 8 B / int32 4 B, the three nlevgrnd patch profiles
    (rootfr read, rresis and rootr written) included; per column its ground
    state (23 FP64 + snl, 7 nlevgrnd profiles read, eff_porosity and
    h2osoi_liqvol written) and per gridcell the forcing (9 FP64 + 1 int32)."""
    from ctsm_amd import abi
    const = _const_fields("ctsm_canopy_state")
    ins = outs = 0
    for name, t, kind in abi.CANOPY_FIELDS:
        if kind not in ("p", "soip"):
            continue
        b = (4 if t == "i32" else 8) * (st.dims.nlevgrnd if kind == "soip" else 1)
        if name in const:
            ins += b
        elif name in _CANOPY_INOUT:
            ins += b
            outs += b
        else:
            outs += b
    NG = st.dims.nlevgrnd
    per_col = 23 * 8 + 4 + NG * 7 * 8 + NG * 2 * 8
    per_grid = 9 * 8 + 4
    return int(st.dims.npatch * (ins + outs) + st.dims.ncol * (per_col + per_grid))


def served_ensemble(ctx, args):
    """Perturbed-parameter members arriving as a Poisson stream, run by the
    continuous-batching server (the library's ctsm_serve_* scheduler) in one
    resident pool; latency = completion - arrival (wall clock), queueing =
    admission - arrival."""
    from ctsm_amd import serve, synthetic
    ctx.set_canopy_params(synthetic.canopy_params(), synthetic.pft_table())
    reqs = serve.make_requests(args.serve_requests, args.serve_rate, seed=0xC75D0004,
                               day_scale=args.serve_day_scale)
    srv, done, wall = serve.serve(ctx, reqs, args.serve_pool_cols, args.serve_pool_cols * 17)
    lat = np.array([q.done - q.arrival for q in done])
    wait = np.array([q.admitted - q.arrival for q in done])
    steps = np.array([q.steps for q in reqs])
    cols = np.array([q.ncol for q in reqs])
    out = {"workload": f"{len(reqs)} members, Poisson {args.serve_rate}/s, 1-20,000 columns (log-uniform, "
                       f"{int(cols.min())}-{int(cols.max())} drawn) with 1-16 patches, 30 parameters x U(0.5,1.5) "
                       f"(soil Clapp-Hornberger and thermal, surface roughness / emissivity / runoff, 18 PFT "
                       f"parameters incl. the stomatal slope on the member's own PFT rows), 1-365 days "
                       f"(log-uniform) x 48 steps x {args.serve_day_scale} ({int(steps.min())}-{int(steps.max())} "
                       f"steps drawn); full land step (CanopyFluxes, SoilTemperature, SoilFluxes, surface water, "
                       f"SoilWater, water tables, drainage) and the energy balance checks (errsoi, errseb, errlon) every step",
           "pool_columns": args.serve_pool_cols, "pool_bytes_per_column": srv.bytes_per_column(),
           "completed": len(done), "wall_s": wall,
           "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
           "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
           "queue_wait_p50_s": float(np.percentile(wait, 50)) if len(wait) else None,
           "queue_wait_p99_s": float(np.percentile(wait, 99)) if len(wait) else None,
           "column_steps_per_s": srv.column_steps / wall,
           "max_abs_errsoi_Wm2": srv.worst_errsoi,
           "max_abs_errseb_Wm2": srv.worst_energy_errors[0],
           "max_abs_errlon_Wm2": srv.worst_energy_errors[1]}
    srv.close()
    return out


def time_events(ctx, fns, reps, prefill=None, inner=1, calls=None):
    """Mean ms of each callable launched `reps` x `inner` times, events on the
    library stream.  For kernels shorter than the host's launch overhead,
    `prefill` enqueues long work first so that the GPU is still busy while the
    host enqueues the timed launches back to back: the events then time the
    kernels, not the host.  `calls`, a list, receives each callable's per-rep
    times."""
    acc = [0.0] * len(fns)
    per = [[] for _ in fns]
    for _ in range(reps):
        if prefill is not None:
            prefill()
        evs = [ctx.record()]
        for fn in fns:
            for _ in range(inner):
                fn()
            evs.append(ctx.record())
        ctx.sync()
        for i in range(len(fns)):
            t = ctx.elapsed_ms(evs[i], evs[i + 1]) / inner
            acc[i] += t
            per[i].append(round(t, 4))
        for e in evs:
            ctx.destroy_event(e)
    if calls is not None:
        calls.extend(per)
    return [a / reps for a in acc]


def _phs_soil_side(ctx, ncol):
    """The two plant-hydraulic-stress soil kernels (root-soil conductances per
    patch, the vertical sink per column) on a synthetic case of ncol columns
    with 0-4 patches each, device-resident, with their HBM rooflines."""
    from ctsm_amd import phs as PH
    A = PH.synthetic(ncol, max_patches=4, seed=3)
    ctx.set_phs_params(PH.default_params())
    D = PH.DevicePhs(ctx, A)
    s = D.struct()
    lay = A.layout()
    nl = A.dims["nlevsoi"]
    cond_ms, sink_ms = time_events(ctx, [lambda: ctx.phs_soil_root_conductance(lay, s, A.npatch, None, dev=True),
                                         lambda: ctx.phs_root_sink(lay, s, A.ncol, None, dev=True)], 3, inner=10)
    D.free()
    # unique bytes: per patch its scalars, rootfr and the three outputs per layer;
    # per column z, dz, hksat, hk_l, smp_l per layer
    cb = A.npatch * (32 + nl * 32) + A.ncol * nl * 40
    used = (A["active"] != 0) & (A["frac_veg_nosno"] > 0) & (A["wtcol"] > 0)
    # per column patch range, z, smp_l and the output per layer, qflx_phs_neg; per
    # patch the three tests and the output, and for contributing patches
    # vegwp(root) and k_soil_root per layer
    sb = A.ncol * (16 + nl * 24) + A.npatch * 24 + int(used.sum()) * (8 + nl * 8)
    roof = lambda b, ms: {"alg_bytes_per_launch": b, "achieved_GBps": b / (ms * 1e-3) / 1e9,
                          "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    out = {"workload": f"{A.ncol:,} columns, {A.npatch:,} patches (0-4 per column), {nl} soil layers",
           "conductance_kernel_ms": cond_ms, "root_sink_kernel_ms": sink_ms,
           "conductance_roofline": dict(kernel="k_phs_conductance", **roof(cb, cond_ms)),
           "root_sink_roofline": dict(kernel="k_phs_root_sink", **roof(sb, sink_ms))}
    # calcstress: the Newton solve of the vegetation water potential, every patch
    # once per call (the input is restored before each timed call)
    B = PH.synthetic_stress(ncol, max_patches=4, seed=4)
    D = PH.DevicePhs(ctx, B)
    s2 = D.struct()
    lay2 = B.layout()
    vw = D.ptr["vegwp"]
    ctx.phs_calcstress(lay2, s2, B.npatch, None, dev=True)
    stress_ms, = time_events(ctx, [lambda: ctx.phs_calcstress(lay2, s2, B.npatch, None, dev=True)], 5,
                             prefill=lambda: ctx.h2d(vw, B["vegwp"]))
    D.free()
    out["calcstress"] = {"patches": B.npatch, "kernel_ms": stress_ms,
                         "patch_solves_per_s": B.npatch / (stress_ms * 1e-3),
                         "bound": "FP64 latency (Newton iterations, 12 pow per iteration)"}
    return out


def _history_and_checkpoint(ctx, ds, structs):
    """The default history tape's per-step update (one launch over every field,
    buffers in HBM) with its HBM roofline, and a checkpoint of the whole
    resident state written and read back through the library."""
    import os
    import shutil
    import tempfile
    from ctsm_amd import restart
    from ctsm_amd.history import HistoryTape
    tape = HistoryTape(ctx, ds)
    elems = sum(f.n1d * f.num2d for f in tape.desc)
    upd_ms, = time_events(ctx, [tape.update], 3, prefill=lambda: ctx.canopyfluxes_dev(ds, structs=structs), inner=10)
    tape.free()
    hb = elems * (8 + 8 + 4 + 8 + 4)        # field read, hbuf and nacs read and written
    out = {"fields": len(tape.specs), "elements": elems, "update_kernel_ms": upd_ms,
           "update_roofline": {"kernel": "k_hist_update", "alg_bytes_per_launch": hb,
                               "achieved_GBps": hb / (upd_ms * 1e-3) / 1e9,
                               "frac": hb / (upd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS}}
    tmp = tempfile.gettempdir()
    need = ds.nbytes()
    if shutil.disk_usage(tmp).free < 3 * need:
        out["checkpoint"] = f"skipped: {need / 1e9:.1f} GB state, too little space in {tmp}"
        return out
    fd, path = tempfile.mkstemp(suffix=".ckpt")
    os.close(fd)
    try:
        t0 = time.perf_counter()
        restart.checkpoint(ctx, ds, path)
        t_write = time.perf_counter() - t0
        nbytes = os.path.getsize(path)
        t0 = time.perf_counter()
        ds2, _ = restart.restore(ctx, path)
        t_read = time.perf_counter() - t0
        ds2.free()
        out.update({"checkpoint_bytes": nbytes, "checkpoint_write_GBps": nbytes / t_write / 1e9,
                    "checkpoint_restore_GBps": nbytes / t_read / 1e9})
    except Exception as e:      # a full or slow file system must not end the benchmark
        out["checkpoint"] = f"failed: {e}"
    finally:
        os.remove(path)
    return out


def extras(ctx, args):
    """The PHS kernels, the history tape's per-step update and a checkpoint
    round trip on a 262,144-column full-column state."""
    from ctsm_amd import synthetic
    from ctsm_amd.state import DeviceState
    st = synthetic.canopy_workload("canopy", ncol=262144)
    ctx.set_canopy_params(synthetic.canopy_params(), synthetic.pft_table())
    ds = DeviceState(st, ctx)
    structs = (*ds.structs(), ds.canopy_struct())
    ctx.land_step_dev(ds, structs=structs)
    out = {"history": _history_and_checkpoint(ctx, ds, structs)}
    ds.free()
    out["phs_soil_side"] = _phs_soil_side(ctx, st.dims.ncol)
    return out


def _const_fields(struct):
    """Names of the `const` (read-only) members of a struct in ctsm_gpu.h."""
    import re
    with open(os.path.join(ROOT, "include", "ctsm_gpu.h")) as f:
        s = f.read()
    body = s[s.index("typedef struct %s {" % struct):s.index("} %s;" % struct)]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = set()
    for line in body.splitlines():
        line = line.strip()
        if line.startswith("const"):
            for nm in line.split(None, 2)[2].rstrip(";").split(","):
                names.add(nm.strip().lstrip("*").strip())
    return names


def surface_water_bytes(st):
    """Algorithmic bytes of one surface water launch: per column 22 FP64 inputs
    (14 of the chain, frac_sno_eff, frac_h2osfc, icefrac and hksat of layers
    1-3), snl, and 16 FP64 outputs (h2osfc included)."""
    return st.dims.ncol * (22 * 8 + 4 + 16 * 8)


def land_step_bytes(st, ncol, npatch):
    """Algorithmic HBM bytes of one full land step per routine: every field a
    routine reads counted once and every field it writes counted once (FP64 8
    B, int32 4 B) over active extents.  The soil solves, CanopyFluxes and
    the surface water chain use the per-kernel counts above; SoilFluxes
    counts its patch members (const = read, others read and written) and its
    column members; the sink reads rootr per patch and writes two profiles
    per column; the water tables and the drainage read their column members
    and liq, ice, dz, z, watsat (and bsw, sucsat, hksat, eff_porosity for the
    drainage) over the soil layers."""
    from ctsm_amd import abi
    NG, NS = st.dims.nlevgrnd, st.dims.nlevsno
    sfc = _const_fields("ctsm_soilflux_state")
    sf_p = sum((8 if t == "f64" else 4) * (1 if n in sfc else 2) for n, t, k in abi.SOILFLUX_FIELDS if k == "p")
    S = -st.col["snl"].astype(np.int64)
    out = {"canopy": canopy_bytes(st), "soiltemp": soiltemp_bytes(st), "richards": richards_bytes(st),
           "surface_water": surface_water_bytes(st),
           "soilfluxes": int(npatch * (sf_p + 8 * 2) + (ncol * (8 * 12) + int((S + 1).sum()) * 8 * 3)),
           "root_sink": int(npatch * (NG * 8 + 8 + 4 + 4) + ncol * (NG * 8 * 2 + 8)),
           "water_table": int(ncol * (7 * 8 + 20 * 8 * 6)),
           "drainage": int(ncol * (15 * 8 + 20 * 8 * 10))}
    return out


def land_cpu_baseline(name, ncol, seconds=15.0, threads=None, per=128):
    """The C oracle's full land step (the same routines as the GPU step) on
    a seeded random sample of the workload's columns, one chunk of `per`
    columns per thread, each thread stepping its chunk until `seconds` pass
    (the oracle runs outside the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    from ctsm_amd import synthetic
    from oracle import oracle as O
    threads = threads or min(16, os.cpu_count() or 1)
    pick = np.random.default_rng(12345).choice(ncol, size=min(ncol, per * threads), replace=False)
    cfg, prm = O.canopy_config(), synthetic.canopy_params()
    pft = O.PftStruct(synthetic.pft_table())
    hydro = synthetic.hydro_params()
    subs = [synthetic.canopy_workload(name, gids=np.sort(pick[i * per:(i + 1) * per])) for i in range(threads)]

    def work(sub):
        n = 0
        t = time.perf_counter()
        while True:
            rc, _ = O.land_step_full(sub, cfg, prm, pft, hydro=hydro, water_table_after=True, drainage_after=True)
            if rc:
                raise RuntimeError(f"oracle land step failed ({rc})")
            n += 1
            if time.perf_counter() - t >= seconds or n >= 48 * 40:
                return n

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        steps = list(ex.map(work, subs))
    el = time.perf_counter() - t0
    total = per * sum(steps)
    return total / el, threads, (f"{threads} threads x {per} columns (a seeded random sample of the "
                                 f"{ncol:,}-column workload, {sum(s.dims.npatch for s in subs):,} patches in all) x "
                                 f"{min(steps)}-{max(steps)} full land steps ({el:.1f} s)")


def headline(ctx, args):
    """The full land step on the config-[3] workload: generation in binned
    order, autotuning, the timed graph replay, per-kernel times of the
    dominant routines, the plain form over the same steps."""
    from ctsm_amd import autotune as AT, synthetic
    from ctsm_amd.state import DeviceState
    t_gen = time.perf_counter()
    ids = synthetic.binned_ids(args.workload, args.ncol, parts=1)
    st = synthetic.canopy_workload(args.workload, gids=ids)
    t_gen = time.perf_counter() - t_gen
    ncol, npatch = st.dims.ncol, st.dims.npatch
    nsno = -st.col["snl"]
    ppc = st.col["patchf"] - st.col["patchi"] + 1
    bytes_by = land_step_bytes(st, ncol, npatch)
    ctx.set_canopy_params(synthetic.canopy_params(), synthetic.pft_table())
    hydro = synthetic.hydro_params()
    ds = DeviceState(st, ctx)
    del st
    full = (*ds.structs(), ds.canopy_struct(), ds.soilflux_struct())
    step = lambda: ctx.land_step_full_dev(ds, structs=full, hydro=hydro, water_table=True, drainage=True)
    for _ in range(args.warmup):
        step()
    ctx.sync()
    res = {"generate_s": round(t_gen, 1), "patches": npatch,
           "mean_snow_layers": float(nsno.mean()), "mean_patches_per_column": float(ppc.mean())}
    if not args.no_autotune:
        split, _, split_table = AT.autotune_canopy(ctx, ds)
        best, _, table = AT.autotune_land(ctx, ds, hydro, steps=2, reps=2)
        res["autotune_canopy_ms"] = {f"{a}/{c}/w{w}": round(ms, 3) for (a, c, w), ms in split_table}
        res["autotune_land_ms_per_step"] = [[c, round(ms, 3)] for c, ms in table]
    if args.set:
        ctx.set_tuning(**{k: int(v) for k, v in (a.split("=", 1) for a in args.set)})
    res["tuning"] = ctx.get_tuning()
    snap = AT.snapshot(ctx, ds)
    # the timed steps, captured once as one graph; the first launch (which
    # uploads the graph) is untimed and the state is restored after it
    graph = ctx.graph_capture_full(ds, nsteps=args.steps, hydro=hydro, water_table=True, drainage=True)
    ctx.graph_launch(graph)
    ctx.sync()
    AT.restore(ctx, ds, snap)
    ctx.sync()
    e0 = ctx.record()
    t0 = time.perf_counter()
    ctx.graph_launch(graph)
    e1 = ctx.record()
    ctx.sync()
    wall = time.perf_counter() - t0
    res["graph_ms_per_step_events"] = ctx.elapsed_ms(e0, e1) / args.steps
    ctx.destroy_event(e0)
    ctx.destroy_event(e1)
    ctx.graph_destroy(graph)
    worst, worst_col = ctx.soil_balance_check_dev(ds)
    nit = np.empty(npatch)
    ctx.d2h(nit, ds.can["num_iter"])
    res["after_steps"] = {"max_abs_errsoi_Wm2": worst, "mean_canopy_iterations_last_step": float(nit.mean()),
                          "max_canopy_iterations_last_step": float(nit.max())}
    # launch times of the step's routines, from the state after the timed day
    # (events on the library stream; each routine launched alone)
    AT.restore(ctx, ds, snap)
    for _ in range(2):
        step()
    structs = (*ds.structs(), ds.canopy_struct())
    k, calls = {}, []
    k["canopy"], = time_events(ctx, [lambda: ctx.canopyfluxes_dev(ds, structs=structs)], 3, calls=calls)
    k["soiltemp"], k["richards"] = time_events(ctx, [lambda: ctx.soiltemperature_dev(ds),
                                                     lambda: ctx.soilwater_dev(ds)], 3, calls=calls)
    res["kernel_ms"] = {n: round(v, 4) for n, v in k.items()}
    # each call: the Richards solve repeated on one sink relaxes the state, so
    # its later calls sub-step less than the solve inside a land step
    res["kernel_ms_calls"] = dict(zip(("canopy", "soiltemp", "richards"), calls))
    res["alg_bytes_per_step"] = bytes_by
    # the plain form over the same steps from the same state
    if not args.no_plain:
        before = ctx.get_tuning()
        ctx.set_tuning(plain_routines=1)
        AT.restore(ctx, ds, snap)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.sync()
        wall_p = time.perf_counter() - t0
        ctx.set_tuning(**before)
        res["plain_ms_per_step"] = wall_p / args.steps * 1e3
        res["plain_column_steps_per_s"] = ncol * args.steps / wall_p
        res["batched_over_plain"] = (ncol * args.steps / wall) / res["plain_column_steps_per_s"]
    AT.free_snapshot(ctx, snap)
    ds.free()
    return wall, res, bytes_by, ncol, npatch


def soil_uneven(ctx, args):
    """The soil step alone (SoilTemperature with SetSoilWaterFractions, then
    SoilWater) on the uneven batch of 65,536 columns, binned by the sub-steps
    of warm-up steps, autotuned (lanes, blocks, 4 streams), 48 steps captured
    as one graph; and its plain form."""
    from ctsm_amd import autotune as AT, synthetic
    from ctsm_amd.state import DeviceState
    n = args.soil_ncol
    st = synthetic.workload("uneven", ncol=n)
    ds = DeviceState(st, ctx)
    for _ in range(max(args.warmup, 1)):
        ctx.step_dev(ds)
    ctx.sync()
    stw = ds.to_host()
    ds.free()
    order = AT.column_order(stw.col["snl"], stw.col["nbedrock"], stw.col["num_substeps"], parts=4) - 1
    st = stw.subset(order)
    before = ctx.get_tuning()
    ds = DeviceState(st, ctx)
    tuning, _, table = AT.autotune(ctx, ds, steps=8, reps=2)
    structs = ds.structs()
    snap = AT.snapshot(ctx, ds)
    graph = ctx.graph_capture(ds, nsteps=args.steps, canopy=False, structs=(*structs, None))
    ctx.graph_launch(graph)
    ctx.sync()
    AT.restore(ctx, ds, snap)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.graph_launch(graph)
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.graph_destroy(graph)
    out = {"workload": f"{n:,} columns, snow 0-12 layers, bedrock 5-20, one patch each; SoilTemperature + "
                       f"SetSoilWaterFractions + SoilWater, {args.steps} steps captured as one graph",
           "column_steps_per_s": n * args.steps / wall, "ms_per_step": wall / args.steps * 1e3,
           "tuning": ctx.get_tuning(), "autotune_ms_per_step": [[c, round(ms, 4)] for c, ms in table]}
    # per-kernel launch times (full-batch launches) and their rooflines
    AT.restore(ctx, ds, snap)
    t_heat, t_rich = time_events(ctx, [lambda: ctx.soiltemperature_dev(ds), lambda: ctx.soilwater_dev(ds)], 8)
    lanes = ctx.get_tuning()["richards_lanes"]
    for nm, ms, b, kn in (("richards", t_rich, richards_bytes(st), ("k_soilwater_g<%d>" % lanes if lanes > 1
                           else "k_soilwater") + " (with its sub-step column order)"),
                          ("soiltemp", t_heat, soiltemp_bytes(st), "k_soiltemp")):
        out[nm] = {"kernel": kn, "ms": ms, "alg_bytes_per_launch": b, "achieved_GBps": b / (ms * 1e-3) / 1e9,
                   "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if not args.no_plain:
        AT.restore(ctx, ds, snap)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.step_dev(ds, plain=True, structs=structs)
        ctx.sync()
        wp = time.perf_counter() - t0
        out["plain_column_steps_per_s"] = n * args.steps / wp
        out["batched_over_plain"] = out["column_steps_per_s"] / out["plain_column_steps_per_s"]
    AT.free_snapshot(ctx, snap)
    ds.free()
    ctx.set_tuning(**before)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["canopy_uneven", "canopy"], default="canopy_uneven")
    ap.add_argument("--ncol", type=int, default=1048576)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--set", action="append", default=[], metavar="NAME=VALUE",
                    help="launch-shape value applied after autotuning (a ctsm_tuning member), e.g. canopy_waves=2")
    ap.add_argument("--soil-ncol", type=int, default=65536,
                    help="columns of the soil-only uneven measurement (0 skips it)")
    ap.add_argument("--extras", action="store_true",
                    help="also time the PHS kernels, the history tape and a checkpoint of a 262,144-column state")
    ap.add_argument("--serve-requests", type=int, default=48,
                    help="served-ensemble measurement: number of requests (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-day-scale", type=float, default=1.0,
                    help="run lengths are log-uniform 1-365 days x 48 steps x this factor")
    ap.add_argument("--serve-pool-cols", type=int, default=131072)
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="run a streaming copy kernel of this size (MB) after the steps, "
                         "to calibrate the HBM byte counters of a profiler run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(columns and their patches shard trivially by contiguous range)"}))
        sys.exit(2)

    from ctsm_amd import lib
    ctx = lib.Context(device=0)
    print(f"[bench] {args.workload} x {args.ncol:,} columns, {args.steps} steps", file=sys.stderr, flush=True)
    wall, res, bytes_by, ncol, npatch = headline(ctx, args)
    value = ncol * args.steps / wall
    ms_step = wall / args.steps * 1e3
    kms = res["kernel_ms"]["canopy"]
    kb = bytes_by["canopy"]
    achieved = kb / (kms * 1e-3) / 1e9
    step_bytes = sum(bytes_by.values())
    res["whole_step_roofline"] = {"alg_bytes_per_step": step_bytes,
                                  "achieved_GBps": step_bytes / (ms_step * 1e-3) / 1e9,
                                  "frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS}
    for nm in ("soiltemp", "richards"):
        ms = res["kernel_ms"][nm]
        kn = "k_soiltemp" if nm == "soiltemp" else "k_soilwater_g"
        p = PROFILED.get((kn, args.workload, ncol))
        if p and any(res["tuning"].get(k) != v for k, v in p["tuning"].items()):
            p = None
        res.setdefault("roofline_other", {})[nm] = {
            "kernel": "k_soiltemp" if nm == "soiltemp" else "k_soilwater_g (with its sub-step column order)",
            "ms": ms, "alg_bytes_per_launch": bytes_by[nm], "achieved_GBps": bytes_by[nm] / (ms * 1e-3) / 1e9,
            "frac": bytes_by[nm] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": p["traffic"] if p else None}
    print(f"[bench] headline {value / 1e6:.2f} M column-steps/s ({ms_step:.2f} ms/step)", file=sys.stderr, flush=True)
    if args.soil_ncol > 0:
        res["soil_only_uneven"] = soil_uneven(ctx, args)
        print("[bench] soil-only batch done", file=sys.stderr, flush=True)
    if args.extras:
        res["extras"] = extras(ctx, args)
    if args.calib_copy_mb > 0:
        nb = args.calib_copy_mb << 20
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy(b, a, nb)
        e0 = ctx.record()
        ctx.copy(b, a, nb)
        e1 = ctx.record()
        ctx.sync()
        ms = ctx.elapsed_ms(e0, e1)
        res["copy_calib"] = {"bytes_read": nb, "bytes_written": nb, "ms": ms, "GBps": 2 * nb / ms / 1e6}
        ctx.free(a)
        ctx.free(b)
    if args.serve_requests > 0:
        res["served_ensemble"] = served_ensemble(ctx, args)
        print("[bench] served ensemble done", file=sys.stderr, flush=True)
    kname = "k_canopy_round"
    prof = PROFILED.get((kname, args.workload, ncol))
    if prof and any(res["tuning"].get(k) != v for k, v in prof["tuning"].items()):
        prof = None
    lanes = res["tuning"]["richards_lanes"]
    out = {
        "metric": "column-steps/sec (FP64, 1 GPU)",
        "value": value,
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, inputs resident in HBM)",
        "config": {"workload": f"{args.workload}: {ncol:,} columns, {npatch:,} patches (1-16 per column), "
                               + ("snow 0-12 layers, bedrock 5-20, " if args.workload == "canopy_uneven" else
                                  "no snow, 20 soil layers, ")
                               + f"full land step (CanopyFluxes, SoilTemperature, SoilFluxes, sink, surface water, "
                                 f"SoilWater, water tables, drainage), dtime 1800 s, {args.steps} steps",
                   "ncol": ncol, "npatch": npatch, "steps": args.steps,
                   "binning": "generated in (snow layers, bedrock) order",
                   "launch": f"the {args.steps} timed steps captured as one graph, launched once",
                   "tuning": ("autotuned" if not args.no_autotune else "default") + "".join(f", {a}" for a in args.set),
                   "richards_lanes": lanes},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": prof["traffic"] if prof else None,
                     "kernel": kname + " (all rounds of one CanopyFluxes call)", "launch_ms": kms,
                     "alg_bytes_per_launch": kb,
                     "fp64_valu_frac": prof["fp64_valu_frac"] if prof else None},
        "detail": res,
    }
    if not args.no_cpu:
        v, threads, sample = land_cpu_baseline(args.workload, ncol, args.cpu_seconds, args.cpu_threads or None)
        out["cpu_baseline"] = {"value": v, "unit": "column-steps/s", "cores": threads, "kind": "port",
                               "sample": sample}
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
