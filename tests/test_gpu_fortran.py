"""SoilWater called from Fortran: a program that holds the column state as
CTSM does (allocatable arrays with the reference's bounds, bound by c_loc) and
calls ctsm_soilwater through the ISO_C_BINDING module
(ctsm_amd/fortran/drive_soilwater.F90).  Its results must equal the oracle's,
bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from ctsm_amd import abi, build, synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,ncol", [("cpu1024", 256), ("uneven", 200)])
def test_soilwater_from_fortran_bitwise(tmp_path, name, ncol):
    exe = build.FORTRAN_DRIVER
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built (no flang)")
    st = S.workload(name, ncol=ncol)
    print(f"case: SoilWater from Fortran, {name}, {ncol} columns", flush=True)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    d = st.dims
    with open(fin, "wb") as f:
        f.write(np.array([d.ncol, d.nlevsno, d.nlevsoi, d.nlevgrnd], dtype=np.int32).tobytes())
        for n, _, _ in abi.COL_FIELDS:
            f.write(np.ascontiguousarray(st.col[n]).tobytes())
    r = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert int(r.stdout.split()[0]) == 0, r.stdout
    raw = fout.read_bytes()
    got, pos = {}, 0
    for n, _, _ in abi.COL_FIELDS:
        a = st.col[n]
        got[n] = np.frombuffer(raw[pos:pos + a.nbytes], dtype=a.dtype).reshape(a.shape)
        pos += a.nbytes
    assert pos == len(raw)
    o = st.copy()
    rc, bad = O.soilwater(o)
    assert rc == 0, bad
    for n in o.col:
        np.testing.assert_array_equal(got[n], o.col[n], err_msg=n)
    assert np.any(o.col["h2osoi_liq"] != st.col["h2osoi_liq"])
