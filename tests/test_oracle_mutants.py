"""Planted mistakes: each entry copies the oracle's C sources, plants one
plausible misreading of the reference (a swapped exponent, a wrong
activation energy, a dropped term, an off-by-one branch), builds the copy and
runs the oracle's pins against it through CTSM_ORACLE_LIB.  Every planted
mistake must make at least one pin fail -- otherwise the pins would let the
same misreading through in the oracle and, since the GPU is held bit-equal
to the oracle, in the kernels.

Nothing here touches the GPU or the product library.
"""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SOURCES = ["ctsm_oracle.c", "canopy_oracle.c", "soilfluxes_oracle.c", "hydrology_oracle.c", "phs_oracle.c",
           "balance_oracle.c"]
PINS = ["tests/test_oracle_balance.py", "tests/test_oracle_watertable.py", "tests/test_oracle_pins.py", "tests/test_oracle_physics.py", "tests/test_oracle_canopy.py",
        "tests/test_oracle_phs_canopy.py"]

# (name, file, text in the oracle, planted text[, occurrences]); the text occurs
# once unless a count is given: the hydraulic-stress photosynthesis restates
# the same kinetics as the plain one (PhotosynthesisMod.F90 has them twice),
# so a misreading of those lines is planted in both.
MUTANTS = [
    ("thk: liquid and ice exponents swapped (SoilTemperatureMod.F90:720)", "ctsm_oracle.c",
     "cm_pow(TKWAT, fl * SOI(S->watsat, c, j)) *\n                    cm_pow(TKICE, (1.0 - fl) * SOI(S->watsat, c, j))",
     "cm_pow(TKWAT, (1.0 - fl) * SOI(S->watsat, c, j)) *\n                    cm_pow(TKICE, fl * SOI(S->watsat, c, j))"),
    ("thk: Kersten number without the +1", "ctsm_oracle.c",
     "dke = fmaxf90(0.0, cm_log10(satw) + 1.0);", "dke = fmaxf90(0.0, cm_log10(satw) + 0.9);"),
    ("thk: frozen Kersten number squared", "ctsm_oracle.c",
     "else              dke = satw;", "else              dke = satw * satw;"),
    ("thk: Kersten branch switch strictly above tfrz", "ctsm_oracle.c",
     "if (T(j) >= TFRZ) dke", "if (T(j) > TFRZ) dke"),
    ("thk: bedrock from layer nbedrock", "ctsm_oracle.c",
     "if (j > nbedrock) th = THK_BEDROCK;", "if (j >= nbedrock) th = THK_BEDROCK;"),
    ("cv: ice and liquid heat capacities swapped", "ctsm_oracle.c",
     "(ICE(j) * CPICE + LIQ(j) * CPLIQ) + 0.0 * CPICE", "(ICE(j) * CPLIQ + LIQ(j) * CPICE) + 0.0 * CPICE"),
    ("cv: solids without the (1 - watsat) factor", "ctsm_oracle.c",
     "A(cv, j) = SOI(S->csol, c, j) * (1.0 - SOI(S->watsat, c, j)) * DZ(j) +",
     "A(cv, j) = SOI(S->csol, c, j) * DZ(j) +"),
    ("tk: interface weights swapped", "ctsm_oracle.c",
     "(A(thk, j) * (Z(j + 1) - ZIv(j)) + A(thk, j + 1) * (ZIv(j) - Z(j)))",
     "(A(thk, j) * (ZIv(j) - Z(j)) + A(thk, j + 1) * (Z(j + 1) - ZIv(j)))"),
    ("fact: top layer without capr", "ctsm_oracle.c",
     "(0.5 * (Z(j) - ZIv(j - 1) + CAPR * (Z(j + 1) - ZIv(j - 1))));",
     "(0.5 * (Z(j) - ZIv(j - 1) + 1.0 * (Z(j + 1) - ZIv(j - 1))));"),
    ("supercool: exponent -b for -1/b", "ctsm_oracle.c",
     "cm_pow(smp / SOI(S->sucsat, c, j), -1.0 / SOI(S->bsw, c, j))",
     "cm_pow(smp / SOI(S->sucsat, c, j), -SOI(S->bsw, c, j))"),
    ("vcmax: jmaxha for vcmaxha (PhotosynthesisMod.F90:1804)", "canopy_oracle.c",
     "PA(S->vcmax_z, p) = vcmax25 * ft(t_veg, P->vcmaxha)", "PA(S->vcmax_z, p) = vcmax25 * ft(t_veg, P->jmaxha)"),
    ("ft: Arrhenius ratio inverted", "canopy_oracle.c",
     "(1.0 - (C_TFRZ + 25.0) / tl)", "(1.0 - tl / (C_TFRZ + 25.0))"),
    ("kc25 from the Ko coefficient", "canopy_oracle.c",
     "const double kc25 = P->kc25_coef * forc_pbot;", "const double kc25 = P->ko25_coef * forc_pbot;", 2),
    ("cp25 without the 0.5", "canopy_oracle.c",
     "const double cp25 = 0.5 * oair / sco;", "const double cp25 = oair / sco;", 2),
    ("lmr: vcmaxha for lmrha", "canopy_oracle.c",
     "lmr25 * ft(t_veg, P->lmrha)", "lmr25 * ft(t_veg, P->vcmaxha)", 2),
    ("vcmax: fth25 normalised with jmaxhd's partner entropy", "canopy_oracle.c",
     "const double vcmaxc = fth25(P->vcmaxhd, vcmaxse);", "const double vcmaxc = fth25(P->vcmaxhd, vcmaxse * 1.01);", 2),
    ("vcmaxse acclimation slope sign", "canopy_oracle.c",
     "const double vcmaxse = (668.39 - 1.07 * t10c)", "const double vcmaxse = (668.39 + 1.07 * t10c)", 2),
    ("C4 low-temperature inhibition sign", "canopy_oracle.c",
     "(1.0 + cm_exp(0.2 * ((C_TFRZ + 15.0) - t_veg)))", "(1.0 + cm_exp(0.2 * (t_veg - (C_TFRZ + 15.0))))", 2),
    ("co-limitation: curvatures swapped", "canopy_oracle.c",
     "if (oracle_quadratic(s->theta_cj, -(s->ac + s->aj)", "if (oracle_quadratic(s->theta_ip, -(s->ac + s->aj)"),
    ("co-limitation: larger root", "canopy_oracle.c",
     "s->ag = fmax90(0.0, fmin90(r1, r2));", "s->ag = fmax90(0.0, fmax90(r1, r2));"),
    ("ZengWang: z0 weights swapped", "canopy_oracle.c",
     "cm_exp(egvf * cm_log(PA(S->z0mv, p)) + (1.0 - egvf) * cm_log(CA(S->z0mg, c)))",
     "cm_exp((1.0 - egvf) * cm_log(PA(S->z0mv, p)) + egvf * cm_log(CA(S->z0mg, c)))"),
    ("ZengWang: egvf with exp(-2 lt)", "canopy_oracle.c",
     "(1.0 - C_ALPHA_AERO * cm_exp(-lt))", "(1.0 - C_ALPHA_AERO * cm_exp(-2.0 * lt))"),
    ("Meier2022: displacement without the square root", "canopy_oracle.c",
     "cm_exp(-cm_pow(C_CD1 * lt, 0.5))", "cm_exp(-(C_CD1 * lt))"),
    ("rb: uaf in place of sqrt(uaf)", "canopy_oracle.c",
     "P->cv / (sqrt(uaf) * sqrt(PA(S->dleaf_patch, p)))", "P->cv / (uaf * sqrt(PA(S->dleaf_patch, p)))"),
    ("under-canopy weight from elai only", "canopy_oracle.c",
     "const double w = cm_exp(-(PA(S->elai, p) + PA(S->esai, p)));", "const double w = cm_exp(-(PA(S->elai, p)));"),
    ("under-canopy wind 0.3 u_m / u*", "canopy_oracle.c",
     "fmin90(0.4, (0.03 * um / ustar))", "fmin90(0.4, (0.3 * um / ustar))"),
    ("QSat: a5 digits transposed", "canopy_oracle.c", "0.196237241e-07", "0.196237421e-07"),
    ("errseb without the canopy heat storage (BalanceCheckMod.F90:989-990)", "soilfluxes_oracle.c",
     "- F->eflx_lh_tot[pp] - F->eflx_soil_grnd[pp] - Cn->dhsdt_canopy[pp];",
     "- F->eflx_lh_tot[pp] - F->eflx_soil_grnd[pp];"),
    ("PHS: getvegwp from the unstressed conductances (PhotosynthesisMod.F90:4051)", "canopy_oracle.c",
     "oracle_phs_getvegwp_one(&s->leaf, x, *gs_sun, *gs_sha, s->qsatl, s->qaf);",
     "oracle_phs_getvegwp_one(&s->leaf, x, gs0sun, gs0sha, s->qsatl, s->qaf);"),
    ("PHS: Ball-Berry minimum conductance without the stress factor (:4431-4432)", "canopy_oracle.c",
     "const double bq = cs[k] * (s->gb_mol - fmax90(bk * s->bbb, 1.0)) - s->mbb * s->an[k] * s->forc_pbot;",
     "const double bq = cs[k] * (s->gb_mol - fmax90(s->bbb, 1.0)) - s->mbb * s->an[k] * s->forc_pbot;"),
    ("PHS: night respiration without bsun (:3516-3520)", "canopy_oracle.c",
     "if (!crop || !modify) *an[k] = PHS2(H->ag_phs, p, k) - *bb[k] * *lmr_z[k];",
     "if (!crop || !modify) *an[k] = PHS2(H->ag_phs, p, k) - *lmr_z[k];"),
    ("PHS: btran weights swapped (:3789-3790)", "canopy_oracle.c",
     "PA(S->btran, p) = *bsun * (laican[SUN] / (laican[SUN] + laican[SHA]))",
     "PA(S->btran, p) = *bsun * (laican[SHA] / (laican[SUN] + laican[SHA]))"),
    ("errseb with the ground heat flux's sign flipped", "soilfluxes_oracle.c",
     "- F->eflx_lh_tot[pp] - F->eflx_soil_grnd[pp] - Cn->dhsdt_canopy[pp];",
     "- F->eflx_lh_tot[pp] + F->eflx_soil_grnd[pp] - Cn->dhsdt_canopy[pp];"),
    ("water mass: top snow layer (j = 0) left out", "balance_oracle.c",
     "for (j = S->snl[cc] + 1; j <= 0; ++j) {", "for (j = S->snl[cc] + 1; j < 0; ++j) {"),
    ("water mass: aquifer counted in hydrologically inactive columns", "balance_oracle.c",
     "if (W->hydrologically_active[cc]) liquid", "if (1) liquid"),
    ("water mass: excess ice left out", "balance_oracle.c",
     "+ (W->excess_ice ? W->excess_ice[cj(L, c, j, 1)] : 0.0);", "+ 0.0;"),
    ("water mass: canopy water of inactive patches counted", "balance_oracle.c",
     "if (Cn->active[pp]) s = s + x[pp] * Cn->wtcol[pp];", "s = s + x[pp] * Cn->wtcol[pp];"),
    ("water mass: canopy water not weighted by wtcol", "balance_oracle.c",
     "if (Cn->active[pp]) s = s + x[pp] * Cn->wtcol[pp];", "if (Cn->active[pp]) s = s + x[pp];"),
    ("water mass: the deepest ground layer left out", "balance_oracle.c",
     "for (j = 1; j <= nlevgrnd; ++j) {", "for (j = 1; j < nlevgrnd; ++j) {"),
    ("errh2o: perched drainage counted as a gain", "balance_oracle.c",
     "- W->qflx_drain[cc] - W->qflx_drain_perched[cc]", "- W->qflx_drain[cc] + W->qflx_drain_perched[cc]"),
    ("errh2o: begwb and endwb swapped", "balance_oracle.c",
     "e = W->endwb[cc] - W->begwb[cc]", "e = W->begwb[cc] - W->endwb[cc]"),
    ("errsol: the diffuse near-infrared band read as the visible one", "balance_oracle.c",
     "+ forc_solai[gg + L->ldg]);", "+ forc_solai[gg]);"),
    ("water table: jwt one layer too deep (:1772)", "ctsm_oracle.c",
     "if (C1(S->zwt, c) <= ZI(S->zi, c, j)) return j - 1;", "if (C1(S->zwt, c) <= ZI(S->zi, c, j)) return j;"),
    ("water table: suction gradient toward the table with the wrong sign (:1801)", "ctsm_oracle.c",
     "const double num = -R->smp[j];", "const double num = R->smp[j];"),
    ("water table: distance to the table from the bottom interface, not the node (:1802)", "ctsm_oracle.c",
     "M_TO_MM * (C1(S->zwt, c) - SNO(S->z, c, j));", "M_TO_MM * (C1(S->zwt, c) - ZI(S->zi, c, j));"),
    ("water table: drainage from the free-drainage expression (:1376-1377)", "ctsm_oracle.c",
     "qcTemp = R.qout[nlayers] + R.dqodw1[nlayers] * dwat[nlayers];",
     "qcTemp = R.hk[nlayers] + R.dhkdw[nlayers] * dwat[nlayers];"),
    ("compute_qcharge: 1.e-3 taken as a double (:2117)", "ctsm_oracle.c",
     "((zwt + (double)1.e-3f) * M_TO_MM)", "((zwt + 1.e-3) * M_TO_MM)"),
    ("compute_qcharge: without the half distance (:2121)", "ctsm_oracle.c",
     "* M_TO_MM * 2.0));", "* M_TO_MM));"),
    ("compute_qcharge: conductivity of layer jwt, not jwt+1 (:2096-2100)", "ctsm_oracle.c",
     "s1 = fmaxf90(R.vwc_liq[jwt + 1] / SOI(S->watsat, c, jwt + 1), 0.01);",
     "s1 = fmaxf90(R.vwc_liq[jw] / SOI(S->watsat, c, jw), 0.01);"),
    ("compute_qcharge: suction of layer jwt+1, not max(1, jwt) (:2103)", "ctsm_oracle.c",
     "smp1 = fmaxf90(C1(S->smpmin, c), R.smp[jw]);", "smp1 = fmaxf90(C1(S->smpmin, c), R.smp[jwt + 1]);"),
    ("compute_qcharge: limit of 1 mm per step (:2125-2126)", "ctsm_oracle.c",
     "q = fmaxf90(-10.0 / dtime, q);\n            q = fminf90(10.0 / dtime, q);",
     "q = fmaxf90(-1.0 / dtime, q);\n            q = fminf90(1.0 / dtime, q);"),
]

CFLAGS = ["-O2", "-mfma", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c99", "-w"]


def _build(tmp, name, old, new, times=1):
    for f in SOURCES:
        shutil.copy(os.path.join(ORACLE, f), tmp)
    path = os.path.join(tmp, name)
    with open(path) as fh:
        s = fh.read()
    assert s.count(old) == times, f"{name}: planted text must occur exactly {times}x: {old!r}"
    with open(path, "w") as fh:
        fh.write(s.replace(old, new))
    lib = os.path.join(tmp, "liboracle.so")
    cmd = ["gcc"] + CFLAGS + ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "ctsm_amd", "csrc"), "-I", ORACLE,
                              "-shared", "-o", lib] + [os.path.join(tmp, f) for f in SOURCES] + ["-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return lib


def _caught(mutant):
    desc, name, old, new, *times = mutant
    with tempfile.TemporaryDirectory() as tmp:
        lib = _build(tmp, name, old, new, *times)
        env = dict(os.environ, CTSM_ORACLE_LIB=lib)
        r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu"]
                           + PINS, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return desc, r.returncode != 0, r.stdout[-400:]


def test_unmutated_copy_passes():
    """The harness itself: the unmodified sources built the same way pass."""
    desc, caught, out = _caught(("identity", "ctsm_oracle.c", "#define TFRZ        273.15",
                                 "#define TFRZ        273.15"))
    assert not caught, out


@pytest.mark.timeout(900)
def test_every_planted_mistake_fails_a_pin():
    assert len(MUTANTS) >= 25
    with ThreadPoolExecutor(max(1, min(6, (os.cpu_count() or 2) - 1))) as ex:
        results = list(ex.map(_caught, MUTANTS))
    survivors = [(d, out) for d, caught, out in results if not caught]
    assert not survivors, "planted mistakes no pin caught:\n" + "\n".join(d for d, _ in survivors)
