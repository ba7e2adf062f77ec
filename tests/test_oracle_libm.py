"""How far the oracle moves when its elementary functions come from the host
C library instead of ctsm_math.h -- CPU only.

The GPU kernels and the oracle share ctsm_math.h, so the GPU parity tests
compare bit for bit.  The reference Fortran calls whatever libm its compiler
ships.  Both are within 1 ulp of the exact result, so the difference between
them is a last-bit choice of pow / log10 / exp / log / atan / sin / tan.
This test runs oracle/liboracle.so and oracle/liboracle_libm.so (the same
sources built with -DORACLE_LIBM) on the same seeded columns of the 1,024
column case, a 4,096-column slice of the uneven batch (snow 0-12, bedrock
5-20) and a 256-column slice of the full-column workload (1-16 patches,
snow and bedrock), for 1 and 48 steps, and bounds the relative difference of
the state north_star names.  It states what a 1e-10 (1 step) / 1e-8 (48
steps) tolerance against a reference on another libm can and cannot mean:
the step is well conditioned (differences stay at 1e-12 or below) except on
two kinds of column.  (a) Columns whose Richards solve sub-steps: the
adaptive schedule halves or accepts a sub-step by comparing an error
estimate with its tolerance, and a last-bit difference that moves the
estimate across the tolerance changes the schedule, so the two runs differ
by the time-stepping error itself (up to ~4e-7 relative after one step, on
the column that takes 35 sub-steps).  (b) Columns holding a layer in mixed
phase at or a few kelvin below freezing: there Phasechange sets the liquid
to the supercooled-water limit watsat (smp / sucsat)**(-1/b),
smp = hfus (tfrz - T) / (grav T), whose relative sensitivity to T is
(1/b) T / (tfrz - T) -- about 40 at 1 K below freezing -- and the latent heat
it releases or takes feeds back into the next step's temperature, so a
last-bit difference grows by 10**2 - 10**4 over a day.  The reference itself
is determined there only to that extent: built on another libm, it differs
from itself by as much.

Each build runs in its own process (the loader binds one library per
process); the parent compares.
"""
import json
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# (case, steps compared): the soil cases run the soil step (SoilTemperature
# with SetSoilWaterFractions, then SoilWater), the canopy case the full land
# step (CanopyFluxes ... drainage).
CASES = ("cpu1024", "uneven4096", "canopy_uneven256")
FIELDS = {"col": ("h2osoi_liq", "h2osoi_ice", "t_soisno", "t_grnd"), "pch": (), "can": ("t_veg",)}

# Bounds on the largest relative difference between the two builds, and on
# the number of columns above north_star's bars.  Measured with gcc 11.4 and
# glibc 2.35 against ctsm_math.h (the report this test prints): after one
# step at most 3.7e-7 (the 1,024-case column with 35 sub-steps; 1 column of
# 1,024 and 4 of 4,096 above 1e-10); after 48 steps at most 4.1e-7, 3
# columns of 1,024 and 3 of 4,096 above 1e-8; the full land step's t_veg
# within 3e-15.  The bounds leave headroom of about 5x.
BOUND_1 = 2e-6       # after one step
BOUND_48 = 1e-5      # after 48 steps
MAX_COLS_1 = 8       # columns above 1e-10 after one step (per case)
MAX_FRAC_48 = 0.01   # fraction of columns above 1e-8 after 48 steps


def _make(case):
    from ctsm_amd import synthetic
    if case == "cpu1024":
        return synthetic.workload("cpu1024")
    if case == "uneven4096":
        return synthetic.workload("uneven", ncol=4096)
    if case == "canopy_uneven256":
        return synthetic.canopy_workload("canopy_uneven", ncol=256)
    raise KeyError(case)


def _run(case, out):
    """Child process: steps the case with whichever oracle CTSM_ORACLE_LIB
    names and saves the compared fields after step 1 and step 48."""
    sys.path.insert(0, ROOT)
    from ctsm_amd import synthetic
    from oracle import oracle as O
    st = _make(case)
    saved = {}
    if case.startswith("canopy"):
        cfg, prm, pft = O.canopy_config(), synthetic.canopy_params(), O.PftStruct(synthetic.pft_table())
        hydro = synthetic.hydro_params()
        step = lambda: O.land_step_full(st, cfg, prm, pft, hydro=hydro, water_table_after=True,
                                        drainage_after=True)
    else:
        step = lambda: O.soil_step(st)
    max_sub = np.zeros(st.dims.ncol)
    for n in range(1, 49):
        rc, bad = step()
        if rc:
            raise SystemExit(f"oracle step {n} failed: rc {rc} at {bad}")
        max_sub = np.maximum(max_sub, st.col["num_substeps"])
        if n in (1, 48):
            for grp, names in FIELDS.items():
                for k in names:
                    if k in getattr(st, grp):
                        saved[f"{n}:{grp}:{k}"] = np.array(getattr(st, grp)[k])
    saved["snl"] = np.array(st.col["snl"])
    saved["nbedrock"] = np.array(st.col["nbedrock"])
    saved["num_substeps"] = np.array(st.col["num_substeps"])
    saved["max_substeps"] = max_sub
    np.savez(out, **saved)


def _child(case, lib, tmp):
    out = os.path.join(tmp, f"{case}_{os.path.basename(lib)}.npz")
    env = dict(os.environ, CTSM_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, os.path.abspath(__file__), case, out], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return dict(np.load(out))


def _reldiff(a, b):
    """Element-wise |a - b| / max(|b|, 1e-12 max|b|), NaN where both are NaN
    (unused snow levels)."""
    ok = np.isfinite(b) | np.isfinite(a)
    a, b = np.where(ok, a, 0.0), np.where(ok, b, 0.0)
    scale = np.maximum(np.abs(b), 1e-12 * max(np.max(np.abs(b)), 1e-300))
    return np.abs(a - b) / scale


@pytest.fixture(scope="module")
def runs():
    libs = [os.path.join(ORACLE, "liboracle.so"), os.path.join(ORACLE, "liboracle_libm.so")]
    for lib in libs:
        if not os.path.exists(lib):
            pytest.skip(f"{lib} is not built (make -C oracle)")
    with tempfile.TemporaryDirectory() as tmp:
        jobs = [(c, lib) for c in CASES for lib in libs]
        with ThreadPoolExecutor(len(jobs)) as ex:
            res = list(ex.map(lambda j: _child(j[0], j[1], tmp), jobs))
    return {c: (res[2 * i], res[2 * i + 1]) for i, c in enumerate(CASES)}


def _report(runs):
    rep = {}
    for case, (a, b) in runs.items():
        for n in (1, 48):
            for key in a:
                if not key.startswith(f"{n}:"):
                    continue
                d = _reldiff(b[key], a[key])
                per_col = d.max(axis=0) if d.ndim == 2 else d
                rep[f"{case} step {n} {key.split(':')[2]}"] = {
                    "max_rel": float(d.max()), "cols_over_1e-10": int((per_col > 1e-10).sum()),
                    "cols_over_1e-8": int((per_col > 1e-8).sum()), "n": int(per_col.size)}
    return rep


def test_libm_sensitivity_is_bounded(runs):
    rep = _report(runs)
    print(json.dumps(rep, indent=1))
    for k, v in rep.items():
        if " step 1 " in k:
            assert v["max_rel"] <= BOUND_1, (k, v)
            assert v["cols_over_1e-10"] <= MAX_COLS_1, (k, v)
        else:
            assert v["max_rel"] <= BOUND_48, (k, v)
            assert v["cols_over_1e-8"] <= MAX_FRAC_48 * v["n"], (k, v)


def test_libm_builds_really_differ(runs):
    """The two builds are different programs: some field moves somewhere
    (otherwise the bounds above would be vacuous)."""
    moved = sum(not np.array_equal(a[k], b[k], equal_nan=True) for a, b in runs.values() for k in a
                if k.startswith("48:"))
    assert moved > 0


def test_sensitive_columns_substep_most_or_hold_mixed_phase(runs):
    """Every column that moves by more than 1e-9 after 48 steps of a soil
    case either moves most on a layer holding both ice and liquid at freezing
    or up to 3 K below it (the supercooled-water limit's sensitivity), or is
    among the columns whose Richards solve takes the most sub-steps (at least
    the batch's 90th percentile of the largest count over the day: the
    adaptive schedule, decided at its error tolerance, is where a last bit
    can change the trajectory); mixed-phase columns occur in each case (so
    the statement is not vacuous)."""
    TFRZ = 273.15
    for case in ("cpu1024", "uneven4096"):
        a, b = runs[case]
        d = np.maximum.reduce([_reldiff(b[f"48:col:{f}"], a[f"48:col:{f}"])
                               for f in ("h2osoi_liq", "h2osoi_ice", "t_soisno")])
        cols = np.nonzero(d.max(axis=0) > 1e-9)[0]
        p90 = np.percentile(a["max_substeps"], 90)
        n_mixed = 0
        for c in cols:
            lev = int(np.argmax(d[:, c]))
            liq, ice, t = (a[f"48:col:{f}"][lev, c] for f in ("h2osoi_liq", "h2osoi_ice", "t_soisno"))
            mixed = liq > 0 and ice > 0 and TFRZ - 3.0 <= t <= TFRZ
            n_mixed += mixed
            assert mixed or a["max_substeps"][c] >= p90, (case, c, lev, liq, ice, t, a["max_substeps"][c], p90)
        assert n_mixed > 0, case


if __name__ == "__main__":
    _run(sys.argv[1], sys.argv[2])
