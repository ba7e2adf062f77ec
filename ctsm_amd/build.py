"""Builds libctsm_mi355x.so (the HIP kernels + C-ABI) in-tree with hipcc for gfx950."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = ["capi.hip", "soilwater.hip", "soiltemp.hip", "plain.hip", "canopy.hip", "soilfluxes.hip", "hydrology.hip",
       "history.hip", "checkpoint.hip", "phs.hip"]
LIB = os.path.join(HERE, "libctsm_mi355x.so")

# -ffp-contract=off: no FMA contraction, every expression rounds as the
#  reference Fortran writes it.
# -packed-fp32-ops off: no packed FP32 instructions anywhere.
# -disable-machine-licm: the machine-level loop-invariant pass hoists every
#  FP64 constant and uniform offset of these long per-column / per-patch
#  loops in front of them, where they exceed the scalar registers and spill
#  into vector-register lanes (a v_readlane per use); left in place they are
#  re-materialised where used.  Measured on the CanopyFluxes kernel: SGPR
#  spills 550 -> 36, scratch 272 -> 80 bytes per lane.
FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-disable-machine-licm",
         "-Wno-unused-command-line-argument"]


def hipcc():
    for p in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if p and (os.path.isabs(p) and os.path.exists(p) or not os.path.isabs(p)):
            return p
    raise RuntimeError("hipcc not found")


OBJDIR = os.path.join(ROOT, "_build")


def build(verbose=False, extra=(), jobs=None):
    """Compiles every source to an object in _build/ (in parallel: each file
    is its own translation unit with its own code object), then links the
    shared library.  Always recompiles."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJDIR, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]
    cflags = [f for f in FLAGS if f != "-shared"] + list(extra)
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in SRC]

    def compile_one(i):
        cmd = [hipcc()] + cflags + inc + ["-c", os.path.join(HERE, "csrc", SRC[i]), "-o", objs[i]]
        if verbose:
            print(" ".join(cmd), flush=True)
        return subprocess.run(cmd, capture_output=True, text=True)

    jobs = jobs or min(len(SRC), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, range(len(SRC))))
    bad = [r for r in results if r.returncode != 0]
    if bad:
        for r in bad:
            sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    # roctx: the trace ranges named after the reference's timers
    cmd = [hipcc()] + FLAGS + list(extra) + objs + ["-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


FORTRAN_DRIVER = os.path.join(HERE, "fortran", "drive_soilwater")
FORTRAN_STEP_DRIVER = os.path.join(HERE, "fortran", "drive_step")
FORTRAN_PHS_DRIVER = os.path.join(HERE, "fortran", "drive_phs")
FORTRAN_STRESS_DRIVER = os.path.join(HERE, "fortran", "drive_stress")


def flang():
    for p in (os.environ.get("FLANG"), "/opt/rocm/lib/llvm/bin/flang"):
        if p and os.path.exists(p):
            return p
    return None


def build_fortran_driver(verbose=False):
    """The Fortran drivers (ctsm_gpu_mod with drive_soilwater.F90, with
    drive_step.F90, with drive_phs.F90 and with drive_stress.F90, linked to the in-tree library);
    returns their paths, or None without flang."""
    fc = flang()
    if fc is None:
        return None
    import tempfile
    fdir = os.path.join(HERE, "fortran")
    for prog, exe in (("drive_soilwater.F90", FORTRAN_DRIVER), ("drive_step.F90", FORTRAN_STEP_DRIVER),
                      ("drive_phs.F90", FORTRAN_PHS_DRIVER), ("drive_stress.F90", FORTRAN_STRESS_DRIVER)):
        with tempfile.TemporaryDirectory() as mod_dir:
            cmd = [fc, "-O2", "-J", mod_dir, os.path.join(fdir, "ctsm_gpu_mod.F90"), os.path.join(fdir, prog),
                   "-L", HERE, "-lctsm_mi355x", "-Wl,-rpath,$ORIGIN/..", "-o", exe + ".tmp"]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("flang failed")
        os.replace(exe + ".tmp", exe)
    return FORTRAN_DRIVER, FORTRAN_STEP_DRIVER, FORTRAN_PHS_DRIVER, FORTRAN_STRESS_DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran_driver(verbose=True))
