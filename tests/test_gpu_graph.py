"""Captured step sequences (ctsm_graph_*) on the GPU: every replay must leave
the state bit-identical to the same steps launched eagerly.

The eager steps run between the replays, with synchronisations, so each
replay is checked after other work on the same stream."""
import numpy as np
import pytest

from ctsm_amd import synthetic as S
from ctsm_amd.state import DeviceState

pytestmark = pytest.mark.gpu


def _same(a, b, groups, tag):
    for grp in groups:
        x, y = getattr(a, grp), getattr(b, grp)
        for k in y:
            np.testing.assert_array_equal(x[k], y[k], err_msg=f"{tag}: {grp}.{k}")


@pytest.mark.parametrize("nsteps", [1, 3])
def test_soil_graph_replays_match_eager(ctx, nsteps):
    st = S.workload("uneven", ncol=2048)
    print(f"case: soil-step graph of {nsteps} steps, 2,048 uneven columns, 3 replays", flush=True)
    de, dg = DeviceState(st, ctx), DeviceState(st, ctx)
    g = None
    try:
        g = ctx.graph_capture(dg, nsteps=nsteps, canopy=False)
        for r in range(3):
            for _ in range(nsteps):
                ctx.step_dev(de)
            ctx.sync()
            ctx.graph_launch(g)
            ctx.sync()
            _same(dg.to_host(), de.to_host(), ("col", "pch"), f"replay {r}")
    finally:
        if g is not None:
            ctx.graph_destroy(g)
        de.free()
        dg.free()


def test_land_graph_replays_match_eager(ctx):
    ctx.set_canopy_params(S.canopy_params(), S.pft_table())
    st = S.canopy_workload("canopy_uneven", ncol=384)
    print(f"case: land-step graph of 2 steps, 384 columns / {st.dims.npatch} patches, 3 replays", flush=True)
    de, dg = DeviceState(st, ctx), DeviceState(st, ctx)
    g = None
    try:
        g = ctx.graph_capture(dg, nsteps=2, canopy=True)
        for r in range(3):
            for _ in range(2):
                ctx.land_step_dev(de)
            ctx.sync()
            ctx.graph_launch(g)
            ctx.sync()
            _same(dg.to_host(), de.to_host(), ("col", "pch", "can"), f"replay {r}")
    finally:
        if g is not None:
            ctx.graph_destroy(g)
        de.free()
        dg.free()


def test_graph_capture_rejects_bad_arguments(ctx):
    st = S.workload("uneven", ncol=64)
    print("case: graph capture argument checks, 64 columns", flush=True)
    ds = DeviceState(st, ctx)
    try:
        from ctsm_amd.lib import CtsmError
        with pytest.raises(CtsmError):
            ctx.graph_capture(ds, nsteps=0, canopy=False)
        with pytest.raises(CtsmError):
            ctx.graph_capture(ds, nsteps=1, canopy=False, nc=ds.dims.ncol + 1)
    finally:
        ds.free()


def test_full_land_step_graph_replays_match_eager():
    """The full land step (SoilFluxes, surface water, water tables, drainage)
    captured for 2 steps; each replay equals 2 eager steps."""
    from ctsm_amd.lib import Context
    from oracle import oracle as O
    c = Context()
    c.set_canopy_params(S.canopy_params(), S.pft_table())
    st = S.canopy_workload("canopy_uneven", ncol=512, seed=0xC75D0081)
    print(f"case: full land-step graph of 2 steps, 512 columns / {st.dims.npatch} patches, 3 replays", flush=True)
    hydro = O.hydro_params()
    de, dg = DeviceState(st, c), DeviceState(st, c)
    g = None
    try:
        g = c.graph_capture_full(dg, nsteps=2, hydro=hydro, water_table=True, drainage=True)
        for r in range(3):
            for _ in range(2):
                c.land_step_full_dev(de, hydro=hydro, water_table=True, drainage=True)
            c.sync()
            c.graph_launch(g)
            c.sync()
            _same(dg.to_host(), de.to_host(), ("col", "pch", "can", "sf", "hy"), f"replay {r}")
    finally:
        if g is not None:
            c.graph_destroy(g)
        de.free()
        dg.free()
        c.close()
