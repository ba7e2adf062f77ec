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
    print("case: ensemble server, 6 members of up to 300 columns, pool 2048 columns", flush=True)
    alone = {}
    for q in reqs:
        serve.build_member(q)
        alone[q.rid] = q.state.copy()
    for q in reqs:
        q.state = None
    try:
        srv, done, wall = serve.serve(ctx, reqs, cap_cols=2048, cap_patches=2048 * 17, max_steps_per_tick=2)
        srv.close()
    finally:
        ctx.close()
    assert len(done) == len(reqs)
    cfg = O.canopy_config()
    pft = O.PftStruct(S.pft_table())
    for q in done:
        o = alone[q.rid]
        for _ in range(q.steps):
            rc, bad = O.land_step(o, cfg, S.canopy_params(), pft)
            assert rc == 0, bad
        for grp in ("col", "pch", "can"):
            a, b = getattr(q.state, grp), getattr(o, grp)
            for k in b:
                np.testing.assert_array_equal(a[k], b[k], err_msg=f"member {q.rid} {grp}.{k}")
