"""Ensemble server on the GPU: members run in a shared pool must end exactly
where the oracle takes each member alone."""
import numpy as np
import pytest

from ctsm_amd import serve, synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_members_in_the_mix_equal_members_alone():
    from ctsm_amd.lib import Context
    ctx = Context()
    ctx.set_canopy_params(S.canopy_params(), S.pft_table())
    reqs = serve.make_requests(6, rate=1e6, seed=21, max_cols=300, max_days=3, steps_per_day=2)
    print("case: ensemble server through the C-ABI scheduler, 6 perturbed members of up to 300 columns, "
          "pool 2048 columns", flush=True)
    alone = {}
    for q in reqs:
        serve.build_member(q)
        alone[q.rid] = (q.state.copy(), q.pft)
    for q in reqs:
        q.state = None
    try:
        srv, done, wall = serve.serve(ctx, reqs, cap_cols=2048, cap_patches=2048 * 17, max_steps_per_tick=2)
        energy = srv.worst_energy_errors
        srv.close()
    finally:
        ctx.close()
    assert len(done) == len(reqs)
    assert srv.worst_errsoi < 1e-7, srv.worst_errsoi
    # errseb / errlon of every patch after every served step, below the reference's warning
    seb, lon = energy
    assert 0.0 <= seb < 1e-7 and 0.0 <= lon < 1e-7, (seb, lon)
    cfg = O.canopy_config()
    hydro = O.hydro_params()
    for q in done:
        o, member_pft = alone[q.rid]
        pft = O.PftStruct(member_pft)          # the member alone, on its own PFT table
        for _ in range(q.steps):
            rc, bad = O.land_step_full(o, cfg, S.canopy_params(), pft, hydro=hydro, water_table_after=True,
                                       drainage_after=True)
            assert rc == 0, bad
        for grp in ("col", "pch", "can", "sf", "hy"):
            a, b = getattr(q.state, grp), getattr(o, grp)
            for k in b:
                np.testing.assert_array_equal(a[k], b[k], err_msg=f"member {q.rid} {grp}.{k}")


def test_served_pool_resumes_from_a_checkpoint(tmp_path):
    from ctsm_amd.lib import Context
    ctx = Context()
    ctx.set_canopy_params(S.canopy_params(), S.pft_table())
    reqs = serve.make_requests(4, rate=1e6, seed=23, max_cols=200, max_days=2, steps_per_day=4)
    print("case: served pool checkpointed after 2 ticks and resumed, 4 members of up to 200 columns", flush=True)
    path = tmp_path / "pool.ckpt"
    try:
        srv = serve.EnsembleServer(ctx, 1024, 1024 * 17, max_steps_per_tick=1)
        for q in reqs:
            serve.build_member(q)
            srv.submit(q)
        for _ in range(3):
            srv.pump(wait=True)
        srv.checkpoint(path)
        srv.drain()
        a = {q.rid: q.state for q in srv.finished}
        srv.close()
        again = {}
        for q in serve.make_requests(4, rate=1e6, seed=23, max_cols=200, max_days=2, steps_per_day=4):
            serve.build_member(q)
            again[q.rid] = q
        srv2 = serve.EnsembleServer.resume(ctx, path, again)
        srv2.drain()
        b = {q.rid: q.state for q in srv2.finished}
        assert srv2.worst_errsoi < 1e-7
        srv2.close()
    finally:
        ctx.close()
    # every member was still running at the checkpoint (4 to 8 steps each, at
    # most 2 enqueued): the resumed pool finishes all of them as the
    # uninterrupted one does
    assert set(a) == set(b) == {q.rid for q in reqs}
    for rid in a:
        for grp in ("col", "pch", "can", "sf", "hy"):
            x, y = getattr(a[rid], grp), getattr(b[rid], grp)
            for k in x:
                np.testing.assert_array_equal(x[k], y[k], err_msg=f"member {rid} {grp}.{k}")
