"""The C-ABI boundary, checked without a GPU.

  - libctsm_mi355x.so loads and exports every function include/ctsm_gpu.h
    declares;
  - the ctypes mirrors (ctsm_amd/abi.py) have the header's sizes and member
    offsets (a probe compiled by gcc against the header prints them);
  - the Fortran ISO_C_BINDING mirrors (ctsm_amd/fortran/ctsm_gpu_mod.F90)
    compile with flang and have the same storage sizes;
  - namelist combinations the library does not run are rejected before any
    device is touched (CTSM_E_CONFIG), and with no GPU ctsm_create reports
    CTSM_E_NODEVICE rather than falling back to anything.
"""
import ctypes as C
import os
import re
import shutil
import subprocess

import pytest

from ctsm_amd import abi, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ctsm_gpu.h")


def _declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ctsm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    declared = _declared_functions()
    assert len(declared) >= 15
    missing = [f for f in declared if not hasattr(L, f)]
    assert not missing, missing
    assert set(declared) == set(lib.EXPORTS)


STRUCTS = (("ctsm_config", abi.Config), ("ctsm_layout", abi.Layout), ("ctsm_tuning", abi.Tuning),
           ("ctsm_col_state", abi.ColState), ("ctsm_patch_state", abi.PatchState),
           ("ctsm_canopy_params", abi.CanopyParams), ("ctsm_pft_params", abi.PftParams),
           ("ctsm_canopy_state", abi.CanopyState), ("ctsm_soilflux_state", abi.SoilFluxState),
           ("ctsm_surface_water_state", abi.SurfaceWaterState), ("ctsm_hydro_params", abi.HydroParams),
           ("ctsm_water_table_state", abi.WaterTableState), ("ctsm_hist_field", abi.HistField),
           ("ctsm_ckpt_field", abi.CkptField), ("ctsm_drainage_state", abi.DrainageState),
           ("ctsm_phs_state", abi.PhsState), ("ctsm_phs_params", abi.PhsParams),
           ("ctsm_phs_stress_state", abi.PhsStressState), ("ctsm_phs_canopy_state", abi.PhsCanopyState),
           ("ctsm_water_balance_state", abi.WaterBalanceState))


def _probe(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ctsm_gpu.h"', 'int main(void){']
    for cname, py in STRUCTS:
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout


def test_struct_layouts_match_header(tmp_path):
    out = _probe(tmp_path)
    table = {}
    for line in out.splitlines():
        parts = line.split()
        table[(parts[0], parts[1])] = int(parts[2])
    for cname, py in STRUCTS:
        assert table[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert table[(cname, f)] == getattr(py, f).offset, (cname, f)


@pytest.mark.skipif(not (shutil.which("flang") or os.path.exists("/opt/rocm/lib/llvm/bin/flang")),
                    reason="flang not available")
def test_fortran_interface_module_matches(tmp_path):
    flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    prog = tmp_path / "sizes.F90"
    prog.write_text("""
program sizes
  use iso_c_binding
  use ctsm_gpu_mod
  type(ctsm_config_t) :: a
  type(ctsm_layout_t) :: b
  type(ctsm_col_state_t) :: c
  type(ctsm_patch_state_t) :: d
  type(ctsm_canopy_params_t) :: e
  type(ctsm_pft_params_t) :: f
  type(ctsm_canopy_state_t) :: g
  type(ctsm_soilflux_state_t) :: h
  type(ctsm_surface_water_state_t) :: i
  type(ctsm_hydro_params_t) :: j
  type(ctsm_water_table_state_t) :: k
  type(ctsm_drainage_state_t) :: l
  type(ctsm_phs_state_t) :: m
  type(ctsm_phs_params_t) :: n
  type(ctsm_phs_stress_state_t) :: o
  type(ctsm_phs_canopy_state_t) :: q
  type(ctsm_water_balance_state_t) :: r
  print '(17I8)', c_sizeof(a), c_sizeof(b), c_sizeof(c), c_sizeof(d), c_sizeof(e), c_sizeof(f), c_sizeof(g), &
       c_sizeof(h), c_sizeof(i), c_sizeof(j), c_sizeof(k), c_sizeof(l), c_sizeof(m), c_sizeof(n), c_sizeof(o), &
       c_sizeof(q), c_sizeof(r)
end program sizes
""")
    mod = os.path.join(ROOT, "ctsm_amd", "fortran", "ctsm_gpu_mod.F90")
    exe = tmp_path / "sizes"
    subprocess.run([flang, "-J", str(tmp_path), mod, str(prog), "-o", str(exe)], check=True, cwd=tmp_path)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert sizes == [C.sizeof(abi.Config), C.sizeof(abi.Layout), C.sizeof(abi.ColState), C.sizeof(abi.PatchState),
                     C.sizeof(abi.CanopyParams), C.sizeof(abi.PftParams), C.sizeof(abi.CanopyState),
                     C.sizeof(abi.SoilFluxState), C.sizeof(abi.SurfaceWaterState), C.sizeof(abi.HydroParams),
                     C.sizeof(abi.WaterTableState), C.sizeof(abi.DrainageState), C.sizeof(abi.PhsState),
                     C.sizeof(abi.PhsParams), C.sizeof(abi.PhsStressState), C.sizeof(abi.PhsCanopyState),
                     C.sizeof(abi.WaterBalanceState)]


def test_trace_ranges_carry_the_reference_timer_names():
    """The library pushes roctx ranges named after the reference's GPTL
    timers around its launches (rocprofv3 --marker-trace --kernel-rename
    reports kernels under them)."""
    out = subprocess.run(["nm", "-D", lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert re.search(r"\bU roctxRangePushA\b", out) and re.search(r"\bU roctxRangePop\b", out)
    data = open(lib.LIB_PATH, "rb").read()
    for name in (b"can_iter", b"canflux", b"SoilTempBandDiag", b"SoilThermProp", b"PhaseChange",
                 b"soiltemperature", b"bgp2", b"hydro_without_drainage", b"hydro2_drainage", b"balchk"):
        assert b"\0" + name + b"\0" in data, name


def test_config_rejected_before_device():
    L = lib.load()
    h = C.c_void_p()
    for bad in (dict(soilwater_movement_method=0), dict(lower_boundary_condition=4),
                dict(upper_boundary_condition=0), dict(nlevsno=13), dict(flux_calculation=42),
                dict(z0param_method=2), dict(stomatalcond_method=0), dict(soil_resis_method=3),
                dict(use_biomass_heat_storage=2), dict(itmax_canopy_fluxes=-1),
                dict(use_hydrstress=2), dict(use_excess_ice=1), dict(modifyphoto_and_lmr_forcrop=2)):
        cfg = lib.default_config(**bad)
        assert L.ctsm_create(C.byref(cfg), 0, C.byref(h)) == lib.CTSM_E_CONFIG, bad


def test_no_silent_fallback_without_gpu():
    """On a machine without a GPU the library refuses to create a context."""
    import torch  # device count only; does not initialise the GPU runtime
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    L = lib.load()
    h = C.c_void_p()
    cfg = lib.default_config()
    assert L.ctsm_create(C.byref(cfg), 0, C.byref(h)) == lib.CTSM_E_NODEVICE
    with pytest.raises(lib.CtsmError):
        lib.Context()
