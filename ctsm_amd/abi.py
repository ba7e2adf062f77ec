"""ctypes mirror of include/ctsm_gpu.h (the library's C-ABI).

The structs are declared once here and shared by the product loader
(ctsm_amd.lib) and the test-side oracle loader.  Field order must match the
header exactly; tests/test_abi.py checks sizes and offsets against the header
by compiling a probe with gcc.
"""
import ctypes as C

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class Config(C.Structure):
    _fields_ = [
        ("nlevsno", C.c_int32),
        ("nlevsoi", C.c_int32),
        ("nlevgrnd", C.c_int32),
        ("soilwater_movement_method", C.c_int32),
        ("upper_boundary_condition", C.c_int32),
        ("lower_boundary_condition", C.c_int32),
        ("flux_calculation", C.c_int32),
        ("snow_thermal_cond_method", C.c_int32),
        ("itmax_canopy_fluxes", C.c_int32),
        ("pad0", C.c_int32),
        ("dtmin", C.c_double),
        ("verySmall", C.c_double),
        ("xTolerUpper", C.c_double),
        ("xTolerLower", C.c_double),
        ("e_ice", C.c_double),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("richards_lanes", C.c_int32),
        ("richards_block", C.c_int32),
        ("soiltemp_block", C.c_int32),
        ("fused_single_kernel", C.c_int32),
    ]


class Layout(C.Structure):
    _fields_ = [
        ("begc_alloc", C.c_int32),
        ("ldc", C.c_int32),
        ("begp_alloc", C.c_int32),
        ("ldp", C.c_int32),
    ]


# (name, ctype, level-kind) in header order.  Level kinds:
#   c   (c)                 p   (p)
#   sno [-nlevsno+1 : nlevgrnd]   zi [-nlevsno : nlevgrnd]
#   soi [1 : nlevgrnd]      snow [-nlevsno+1 : 0]   lyr [-nlevsno+1 : 1] (patch)
COL_FIELDS = [
    ("snl", "i32", "c"),
    ("nbedrock", "i32", "c"),
    ("patchi", "i32", "c"),
    ("patchf", "i32", "c"),
    ("dz", "f64", "sno"),
    ("z", "f64", "sno"),
    ("zi", "f64", "zi"),
    ("watsat", "f64", "soi"),
    ("bsw", "f64", "soi"),
    ("sucsat", "f64", "soi"),
    ("hksat", "f64", "soi"),
    ("tkmg", "f64", "soi"),
    ("tkdry", "f64", "soi"),
    ("csol", "f64", "soi"),
    ("eff_porosity", "f64", "soi"),
    ("smp_l", "f64", "soi"),
    ("hk_l", "f64", "soi"),
    ("thk", "f64", "sno"),
    ("icefrac", "f64", "soi"),
    ("qcharge", "f64", "c"),
    ("num_substeps", "f64", "c"),
    ("qin", "f64", "soi"),
    ("qout", "f64", "soi"),
    ("h2osoi_liq", "f64", "sno"),
    ("h2osoi_ice", "f64", "sno"),
    ("h2osfc", "f64", "c"),
    ("h2osno_no_layers", "f64", "c"),
    ("int_snow", "f64", "c"),
    ("frac_sno_eff", "f64", "c"),
    ("frac_h2osfc", "f64", "c"),
    ("snow_depth", "f64", "c"),
    ("bw", "f64", "snow"),
    ("snomelt_accum", "f64", "c"),
    ("qflx_infl", "f64", "c"),
    ("qflx_rootsoi", "f64", "soi"),
    ("qflx_snomelt", "f64", "c"),
    ("qflx_snofrz", "f64", "c"),
    ("qflx_snow_drain", "f64", "c"),
    ("qflx_snomelt_lyr", "f64", "snow"),
    ("qflx_snofrz_lyr", "f64", "snow"),
    ("qflx_h2osfc_to_ice", "f64", "c"),
    ("t_soisno", "f64", "sno"),
    ("t_h2osfc", "f64", "c"),
    ("t_grnd", "f64", "c"),
    ("fact", "f64", "sno"),
    ("xmf", "f64", "c"),
    ("xmf_h2osfc", "f64", "c"),
    ("c_h2osfc", "f64", "c"),
    ("imelt", "i32", "sno"),
    ("emg", "f64", "c"),
    ("htvp", "f64", "c"),
    ("eflx_bot", "f64", "c"),
    ("eflx_fgr12", "f64", "c"),
    ("eflx_fgr", "f64", "soi"),
    ("eflx_snomelt", "f64", "c"),
    ("eflx_h2osfc_to_snow", "f64", "c"),
    ("forc_lwrad", "f64", "c"),
]

PATCH_FIELDS = [
    ("column", "i32", "p"),
    ("wtcol", "f64", "p"),
    ("frac_veg_nosno", "i32", "p"),
    ("sabg", "f64", "p"),
    ("sabg_soil", "f64", "p"),
    ("sabg_snow", "f64", "p"),
    ("sabg_lyr", "f64", "lyr"),
    ("dlrad", "f64", "p"),
    ("cgrnd", "f64", "p"),
    ("eflx_sh_grnd", "f64", "p"),
    ("eflx_sh_snow", "f64", "p"),
    ("eflx_sh_soil", "f64", "p"),
    ("eflx_sh_h2osfc", "f64", "p"),
    ("qflx_evap_soi", "f64", "p"),
    ("qflx_ev_snow", "f64", "p"),
    ("qflx_ev_soil", "f64", "p"),
    ("qflx_ev_h2osfc", "f64", "p"),
    ("eflx_gnet", "f64", "p"),
    ("dgnetdT", "f64", "p"),
    ("sabg_chk", "f64", "p"),
]

_CT = {"i32": i32p, "f64": f64p}


class ColState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in COL_FIELDS]


class PatchState(C.Structure):
    _fields_ = [(n, _CT[t]) for n, t, _ in PATCH_FIELDS]


def nlev_of(kind, nlevsno, nlevgrnd):
    """Number of levels stored for a level kind (1 for scalars)."""
    return {
        "c": 1,
        "p": 1,
        "sno": nlevsno + nlevgrnd,
        "zi": nlevsno + nlevgrnd + 1,
        "soi": nlevgrnd,
        "snow": nlevsno,
        "lyr": nlevsno + 1,
    }[kind]


def lo_of(kind, nlevsno):
    """Fortran lower level bound of a level kind."""
    return {"c": 0, "p": 0, "sno": 1 - nlevsno, "zi": -nlevsno, "soi": 1,
            "snow": 1 - nlevsno, "lyr": 1 - nlevsno}[kind]
