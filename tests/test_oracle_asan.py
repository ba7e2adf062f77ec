"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer -- CPU only.

oracle/_asan/liboracle.so (`make -C oracle asan`: the same sources at -O1
with -fsanitize=address,undefined and no recovery) is loaded in a child
process with the sanitizer runtime preloaded and driven through every entry
point the tests use: the soil step on the uneven batch (snow 0-12, bedrock
5-20), the full land step with surface water, water table and drainage on
the 1-16 patch workload under a patch filter that skips patches, the energy
and soil balance checks, and the plant hydraulic stress routines.  A stray
read or write past an array, a use after free or undefined arithmetic
(signed overflow, a shift out of range, a misaligned load) aborts the child
with the sanitizer's report, and the test fails with it.  The results must
also equal the regular build's (the sanitizers change nothing they do not
report)."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
ASAN_LIB = os.path.join(ORACLE, "_asan", "liboracle.so")


def _runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def _drive(out):
    """Child process: every oracle entry point on small cases; saves what
    they wrote."""
    sys.path.insert(0, ROOT)
    from ctsm_amd import phs as PH, synthetic as S
    from oracle import oracle as O
    res = {}
    st = S.workload("uneven", ncol=256)
    for _ in range(3):
        rc, bad = O.soil_step(st)
        assert rc == 0, (rc, bad)
    res["soil_t"] = st.col["t_soisno"]
    res["soil_liq"] = st.col["h2osoi_liq"]
    st = S.canopy_workload("canopy_uneven", ncol=48)
    cfg, prm, pft = O.canopy_config(), S.canopy_params(), O.PftStruct(S.pft_table())
    for _ in range(2):
        rc, bad = O.land_step_full(st, cfg, prm, pft, hydro=S.hydro_params(), water_table_after=True,
                                   drainage_after=True)
        assert rc == 0, (rc, bad)
    filt = np.arange(1, st.dims.npatch + 1, 2, dtype=np.int32)
    rc, bad = O.canopyfluxes(st, cfg, prm, pft, filter_p=filt)
    assert rc == 0, (rc, bad)
    res["t_veg"] = st.can["t_veg"]
    res["energy"] = np.array(O.energy_balance(st)[:2], dtype=float)
    res["soilbal"] = np.array(O.soil_balance(st)[:1], dtype=float)
    A = PH.synthetic(64, seed=3)
    assert O.phs_soil_root_conductance(A, PH.default_params()) == 0
    assert O.phs_root_sink(A) == 0
    B = PH.synthetic_stress(32, seed=4)
    assert O.phs_calcstress(B, PH.default_params()) == 0
    for k in ("k_soil_root", "root_conductance", "qflx_rootsoi", "qflx_hydr_redist"):
        res["phs_" + k] = A[k]
    for k in ("vegwp", "bsun", "bsha", "qflx_tran_veg"):
        res["stress_" + k] = B[k]
    np.savez(out, **{k: np.asarray(v, dtype=float) for k, v in res.items()})


def _child(tmp_path, lib, preload):
    out = str(tmp_path / (os.path.basename(os.path.dirname(lib)) + ".npz"))
    env = dict(os.environ, CTSM_ORACLE_LIB=lib)
    if preload:
        env["LD_PRELOAD"] = preload
        # the interpreter's own allocations are not the oracle's to answer for
        env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
        env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    r = subprocess.run([sys.executable, os.path.abspath(__file__), out], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=900)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, \
        r.stdout[-4000:] + r.stderr[-8000:]
    return dict(np.load(out))


def test_oracle_is_clean_under_address_and_ub_sanitizers(tmp_path):
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not (shutil.which("make") and asan and ubsan):
        pytest.skip("gcc's sanitizer runtimes are not installed")
    subprocess.run(["make", "-s", "-C", ORACLE, "asan", "all"], check=True)
    got = _child(tmp_path, ASAN_LIB, f"{asan}:{ubsan}")
    ref = _child(tmp_path, os.path.join(ORACLE, "liboracle.so"), None)
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


if __name__ == "__main__":
    _drive(sys.argv[1])
