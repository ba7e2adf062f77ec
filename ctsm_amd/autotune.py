"""Launch-shape autotuner and column binning for the soil step.

Two decisions shape the step on an MI355X, and neither changes a single bit
of the results (columns are independent; tests/test_gpu_parity.py checks every
shape against the oracle):

  * column order: the Richards sub-step count of a column varies from 1 to
    ~35 trials and a wavefront runs as long as its slowest lane, so columns
    are ordered by (bedrock depth, sub-steps taken in the previous step,
    snow layers) -- the previous step is a good predictor of the next;
  * launch shape: lanes per column in the Richards solve (1, 2, 4, 8), block
    sizes, one fused kernel vs. heat-solve kernel + Richards kernel, and the
    number of concurrent streams the step's columns are split across.

autotune() times each candidate shape for a few steps on a snapshot of the
resident state (restored between candidates with device copies) and keeps
the fastest on the context.
"""
import numpy as np

DEFAULT_CANDIDATES = [
    dict(richards_lanes=8, richards_block=128, soiltemp_block=64, fused_single_kernel=0, step_streams=4),
    dict(richards_lanes=8, richards_block=128, soiltemp_block=128, fused_single_kernel=0, step_streams=4),
    dict(richards_lanes=0, richards_block=128, soiltemp_block=64, fused_single_kernel=0, step_streams=4),
    dict(richards_lanes=8, richards_block=128, soiltemp_block=64, fused_single_kernel=0, step_streams=2),
    dict(richards_lanes=8, richards_block=128, soiltemp_block=64, fused_single_kernel=0, step_streams=1),
    dict(richards_lanes=8, richards_block=256, soiltemp_block=64, fused_single_kernel=0, step_streams=4),
    dict(richards_lanes=4, richards_block=128, soiltemp_block=128, fused_single_kernel=0, step_streams=4),
    dict(richards_lanes=2, richards_block=128, soiltemp_block=128, fused_single_kernel=0, step_streams=4),
    dict(richards_lanes=1, richards_block=64, soiltemp_block=128, fused_single_kernel=0, step_streams=4),
    dict(richards_lanes=1, richards_block=64, soiltemp_block=64, fused_single_kernel=1, step_streams=1),
]


def column_order(snl, nbedrock, prev_substeps=None, parts=1):
    """1-based column order that groups similar work into wavefronts.

    parts > 1 deals the sorted columns over that many contiguous parts
    (part k takes every parts-th column of the sorted order), so a step split
    into contiguous parts across streams gives every part the same work mix
    while each part stays sorted."""
    snl = np.asarray(snl)
    nb = np.asarray(nbedrock)
    if prev_substeps is None:
        keys = (snl, nb)
    else:
        keys = (snl, np.asarray(prev_substeps), nb)
    o = np.lexsort(keys)
    if parts > 1:
        o = np.concatenate([o[k::parts] for k in range(parts)])
    return (o + 1).astype(np.int32)


def snapshot(ctx, ds):
    """Device copy of every field of a DeviceState (returns dict of pointers)."""
    snap = {}
    for group, d in (("c", ds.col), ("p", ds.pch), ("g", ds.can), ("s", getattr(ds, "sf", {})),
                     ("y", getattr(ds, "hy", {}))):
        for k, ptr in d.items():
            shape, dt = ds._shapes[(group, k)]
            nb = int(np.prod(shape)) * np.dtype(dt).itemsize
            q = ctx.alloc(nb)
            ctx.copy(q, ptr, nb)
            snap[(group, k)] = (q, nb)
    ctx.sync()
    return snap


def restore(ctx, ds, snap):
    for (group, k), (q, nb) in snap.items():
        dst = {"c": ds.col, "p": ds.pch, "g": ds.can, "s": getattr(ds, "sf", {}), "y": getattr(ds, "hy", {})}[group][k]
        ctx.copy(dst, q, nb)


def free_snapshot(ctx, snap):
    ctx.sync()
    for q, _ in snap.values():
        ctx.free(q)


def autotune(ctx, ds, filter_dev=None, n=None, steps=8, reps=2, candidates=None, dtime=1800.0, verbose=False):
    """Pick the fastest launch shape for this state, timing `steps` steps in
    one multi-step call (as a batched run issues them), best of `reps` runs
    from the same snapshot.  Leaves `ds` unchanged."""
    candidates = candidates or DEFAULT_CANDIDATES
    snap = snapshot(ctx, ds)
    structs = ds.structs()
    best, best_ms, table = None, float("inf"), []
    for cand in candidates:
        ctx.set_tuning(**cand)
        restore(ctx, ds, snap)
        ctx.step_dev(ds, filter_dev=filter_dev, n=n, dtime=dtime, structs=structs)   # warm
        ctx.sync()
        ms = float("inf")
        for _ in range(reps):
            restore(ctx, ds, snap)
            e0 = ctx.record()
            ctx.steps_dev(ds, steps, filter_dev=filter_dev, n=n, dtime=dtime, structs=structs)
            e1 = ctx.record()
            ctx.sync()
            ms = min(ms, ctx.elapsed_ms(e0, e1) / steps)
            ctx.destroy_event(e0)
            ctx.destroy_event(e1)
        table.append((dict(cand), ms))
        if verbose:
            print(f"autotune {cand} -> {ms:.3f} ms/step", flush=True)
        if ms < best_ms:
            best, best_ms = cand, ms
    restore(ctx, ds, snap)
    free_snapshot(ctx, snap)
    ctx.set_tuning(**best)
    return best, best_ms, table


# (first-round iterations, later-round iterations) of the CanopyFluxes
# leaf-temperature iteration; first > itmax is the one-pass kernel.
CANOPY_CANDIDATES = [(1001, 1), (4, 2), (4, 4), (5, 3), (6, 3), (6, 6), (8, 4), (10, 5)]
# register budgets (waves per SIMD) of the CanopyFluxes kernel
CANOPY_WAVES = (1, 2, 3)


def autotune_canopy(ctx, ds, filter_dev=None, n=None, reps=2, candidates=None, waves=None, dtime=1800.0,
                    verbose=False):
    """Pick the CanopyFluxes launch shape that runs this state fastest (timed
    on a snapshot of the resident state, which is left unchanged): the round
    split over `candidates` at the current register budget, then the budget
    over `waves` at the best split.  Returns ((first, later, waves), ms,
    table of ((first, later, waves), ms))."""
    candidates = candidates or CANOPY_CANDIDATES
    waves = waves or CANOPY_WAVES
    snap = snapshot(ctx, ds)
    structs = (*ds.structs(), ds.canopy_struct())
    best, best_ms, table = None, float("inf"), []

    def timed(first, later, w):
        ctx.set_tuning(canopy_first_iterations=first, canopy_round_iterations=later, canopy_waves=w)
        ms = 0.0
        for _ in range(reps):
            restore(ctx, ds, snap)
            e0 = ctx.record()
            ctx.canopyfluxes_dev(ds, filterp_dev=filter_dev, n=n, dtime=dtime, structs=structs)
            e1 = ctx.record()
            ctx.sync()
            ms += ctx.elapsed_ms(e0, e1) / reps
            ctx.destroy_event(e0)
            ctx.destroy_event(e1)
        table.append(((first, later, w), ms))
        if verbose:
            print(f"autotune canopy first={first} later={later} waves={w} -> {ms:.3f} ms", flush=True)
        return ms

    w0 = ctx.get_tuning()["canopy_waves"]
    for first, later in candidates:
        ms = timed(first, later, w0)
        if ms < best_ms:
            best, best_ms = (first, later, w0), ms
    for w in waves:
        if w == w0:
            continue
        ms = timed(best[0], best[1], w)
        if ms < best_ms:
            best, best_ms = (best[0], best[1], w), ms
    restore(ctx, ds, snap)
    free_snapshot(ctx, snap)
    ctx.set_tuning(canopy_first_iterations=best[0], canopy_round_iterations=best[1], canopy_waves=best[2])
    return best, best_ms, table


# Launch shapes of the soil solves inside a full land step (one stream, so
# step_streams does not apply): lanes per column and block sizes.
LAND_CANDIDATES = [
    dict(richards_lanes=0, richards_block=128, soiltemp_block=64),
    dict(richards_lanes=8, richards_block=128, soiltemp_block=64),
    dict(richards_lanes=8, richards_block=128, soiltemp_block=128),
    dict(richards_lanes=8, richards_block=256, soiltemp_block=64),
    dict(richards_lanes=8, richards_block=64, soiltemp_block=64),
    dict(richards_lanes=4, richards_block=128, soiltemp_block=64),
    dict(richards_lanes=4, richards_block=256, soiltemp_block=64),
    dict(richards_lanes=2, richards_block=128, soiltemp_block=64),
    dict(richards_lanes=1, richards_block=64, soiltemp_block=64),
]


def autotune_land(ctx, ds, hydro, steps=2, reps=2, candidates=None, dtime=1800.0, verbose=False):
    """Pick the soil launch shape that runs `steps` full land steps (surface
    water, water tables and drainage included) fastest on this state, best of
    `reps` runs from the same snapshot; the state is left unchanged."""
    candidates = candidates or LAND_CANDIDATES
    snap = snapshot(ctx, ds)
    full = (*ds.structs(), ds.canopy_struct(), ds.soilflux_struct())
    best, best_ms, table = None, float("inf"), []
    for cand in candidates:
        ctx.set_tuning(**cand)
        ms = float("inf")
        for _ in range(reps):
            restore(ctx, ds, snap)
            e0 = ctx.record()
            for _ in range(steps):
                ctx.land_step_full_dev(ds, structs=full, hydro=hydro, water_table=True, drainage=True, dtime=dtime)
            e1 = ctx.record()
            ctx.sync()
            ms = min(ms, ctx.elapsed_ms(e0, e1) / steps)
            ctx.destroy_event(e0)
            ctx.destroy_event(e1)
        table.append((dict(cand), ms))
        if verbose:
            print(f"autotune land {cand} -> {ms:.3f} ms/step", flush=True)
        if ms < best_ms:
            best, best_ms = cand, ms
    restore(ctx, ds, snap)
    free_snapshot(ctx, snap)
    ctx.set_tuning(**best)
    return best, best_ms, table
