"""bench.py's host side on the CPU: the module imports, every autotuner it
calls exists, the algorithmic byte counts of a small workload of the
headline kind are positive and add up, and --gpus above 1 is refused."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_helpers_and_autotuners():
    sys.path.insert(0, ROOT)
    import bench
    from ctsm_amd import autotune as AT, synthetic
    for name in ("autotune", "autotune_canopy", "autotune_land", "snapshot", "restore", "free_snapshot",
                 "column_order"):
        assert callable(getattr(AT, name)), name
    ids = synthetic.binned_ids("canopy_uneven", 256, parts=1)
    assert sorted(ids.tolist()) == list(range(256))
    st = synthetic.canopy_workload("canopy_uneven", gids=ids)
    b = bench.land_step_bytes(st, st.dims.ncol, st.dims.npatch)
    assert set(b) >= {"canopy", "soiltemp", "richards", "soilfluxes", "root_sink"}
    assert all(v > 0 for v in b.values())
    # the Richards count is 120 L + 28 per column, L = the bedrock index
    assert b["richards"] == int(np.sum(120 * st.col["nbedrock"].astype(np.int64) + 28))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2 and "one GPU" in r.stdout
