"""Pins of the plant-hydraulic-stress oracle (oracle/phs_oracle.c).

  * plc / d1plc against the reference's own known answers
    (src/biogeophys/test/Photosynthesis_test/test_Photosynthesis.pf:65-70 with
    setParamsForTesting, PhotosynthesisMod.F90:930-931: pft 1, psi50 = -150000
    mm, ck = 3.95, x = -1000 * nlevgrnd = -25000 mm, tolerance 1e-13), plc(0) = 1,
    plc(psi50) = 1/2 exactly, the 0.005 cut, d1plc against a central difference;
  * the root-soil conductances against the Gardner closed form (a root length
    density of 1/pi gives r_soil = 1 m, so k_soil = min(hksat, hk_l) / 1000), and
    against a numpy evaluation in a different algebraic form (the series law as
    k_s k_r / (k_s + k_r)); the top layer and rootless layers at zero, the 1 gC/m2
    biomass floor, the 1e-16 floors past the vulnerability cut;
  * the vertical sink against a numpy sum over patches, zero exchange at
    hydrostatic equilibrium (smp = vegwp(root) + 1000 z, exact in binary64),
    and inactive, snow-buried and zero-weight patches left out.
"""
import numpy as np
import pytest

from ctsm_amd import phs as PH
from oracle import oracle as O

PSI50, CK = -150000.0, 3.95


def test_plc_known_answers_of_the_reference_test():
    # test_Photosynthesis.pf:67 and :70
    assert abs(O.phs_plc(-25000.0, PSI50, CK) - 0.999415208562283) <= 1e-13
    assert abs(O.phs_d1plc(-25000.0, PSI50, CK) - 9.237002539040872e-8) <= 1e-13


def test_plc_special_values_and_cut():
    assert O.phs_plc(0.0, PSI50, CK) == 1.0
    assert O.phs_plc(PSI50, PSI50, CK) == 0.5
    assert O.phs_plc(1.6 * PSI50, PSI50, CK) == pytest.approx(2.0 ** -(1.6 ** CK), rel=1e-14)
    assert O.phs_plc(1.6 * PSI50, PSI50, CK) > 0.005
    assert O.phs_plc(1.7 * PSI50, PSI50, CK) == 0.0          # 2**-(1.7**3.95) < 0.005


@pytest.mark.parametrize("x", [-2000.0, -60000.0, -150000.0, -210000.0])
def test_d1plc_is_the_derivative_of_plc(x):
    h = 1e-4 * abs(x)
    fd = (O.phs_plc(x + h, PSI50, CK) - O.phs_plc(x - h, PSI50, CK)) / (2 * h)
    assert O.phs_d1plc(x, PSI50, CK) == pytest.approx(fd, rel=1e-6)


def _single(nl=20):
    A = PH.PhsArrays(1, 1)
    A["patchi"][:] = 1
    A["patchf"][:] = 1
    A["column"][:] = 1
    A["active"][:] = 1
    A["frac_veg_nosno"][:] = 1
    A["wtcol"][:] = 1.0
    A["dz"][12:, 0] = 0.1
    A["z"][12:, 0] = 0.05 + 0.1 * np.arange(25)
    A["hksat"][:] = 1e-2
    A["hk_l"][:] = 1e-3
    A["smp_l"][:] = -1000.0
    A["tsai"][:] = 1.0
    A["tlai"][:] = 3.0
    A["rootfr"][:nl, 0] = 1.0 / nl
    return A


# the four segments (sun, sha, xyl, root) get distinct curves; the root one is PSI50, CK
SEG_PSI50 = np.array([[-90000.0], [-120000.0], [-400000.0], [PSI50]])
SEG_CK = np.array([[2.0], [2.5], [6.0], [CK]])


def _params(root_density=0.31e6, radius=0.29e-3, krmax=5e-9, froot_leaf=1.5):
    return PH.phs_params([krmax], [radius], [root_density], [froot_leaf], SEG_PSI50, SEG_CK)


def test_gardner_closed_form_at_unit_soil_radius():
    # froot_carbon = 0: the biomass floor 2 g/m3; root_density = 2 / r**2 makes
    # the root length density 1/pi and r_soil = 1 m
    A = _single()
    r = 0.29e-3
    assert O.phs_soil_root_conductance(A, _params(root_density=2.0 / r ** 2, radius=r)) == 0
    ks = A["soil_conductance"][:20, 0]
    assert np.allclose(ks, np.minimum(1e-2, 1e-3) / 1000.0, rtol=1e-14, atol=0)
    A["hk_l"][:] = 5e-2                                     # hksat the smaller one now
    O.phs_soil_root_conductance(A, _params(root_density=2.0 / r ** 2, radius=r))
    assert np.allclose(A["soil_conductance"][:20, 0], 1e-5, rtol=1e-14, atol=0)


def _numpy_conductance(A, krmax, radius, density, froot_leaf, psi50, ck):
    nl = 20
    p, c = 0, 0
    fr = A["rootfr"][:nl, p]
    dz = A["dz"][12:12 + nl, c]
    z = A["z"][12:12 + nl, c]
    b = np.maximum(2.0 * A["froot_carbon"][p] * fr / dz, 2.0)
    L = b / (density * np.pi * radius ** 2)
    ks = np.maximum(np.minimum(A["hksat"][:nl, c], A["hk_l"][:nl, c]) * np.sqrt(np.pi * L) / 1000.0, 1e-16)
    plc = 2.0 ** -((A["smp_l"][:nl, c] / psi50) ** ck)
    plc[plc < 0.005] = 0.0
    rai = (A["tsai"][p] + A["tlai"][p]) * froot_leaf * fr
    kr = np.maximum(plc * rai * krmax / (0.25 + z), 1e-16)
    k = ks * kr / (ks + kr)
    k[(rai * fr <= 0) | (np.arange(1, nl + 1) == 1)] = 0.0
    return ks, kr, k


@pytest.mark.parametrize("froot", [0.0, 40.0, 900.0])
def test_conductances_against_numpy(froot):
    A = _single(nl=14)
    A["froot_carbon"][:] = froot
    A["smp_l"][:14, 0] = -np.geomspace(50.0, 6e5, 14)           # the driest rooted layers past the cut
    A["hk_l"][:, 0] = np.geomspace(1e-9, 5e-2, 25)
    assert O.phs_soil_root_conductance(A, _params()) == 0
    ks, kr, k = _numpy_conductance(A, 5e-9, 0.29e-3, 0.31e6, 1.5, PSI50, CK)
    assert np.allclose(A["soil_conductance"][:20, 0], ks, rtol=1e-12, atol=0)
    assert np.allclose(A["root_conductance"][:20, 0], kr, rtol=1e-12, atol=0)
    assert np.allclose(A["k_soil_root"][:20, 0], k, rtol=1e-12, atol=0)
    assert A["k_soil_root"][0, 0] == 0.0                        # top layer
    assert np.all(A["k_soil_root"][14:20, 0] == 0.0)            # no roots
    assert np.all(A["root_conductance"][14:20, 0] == 1e-16)
    cut = A["smp_l"][:14, 0] < 1.7 * PSI50
    assert cut.any() and np.all(A["root_conductance"][:14, 0][cut] == 1e-16)
    assert np.all(A["root_conductance"][:14, 0][~cut] > 1e-16)


def test_bad_type_is_reported():
    A = _single()
    A["itype"][:] = 3                                           # only one PFT in the table
    assert O.phs_soil_root_conductance(A, _params()) == 1


def _numpy_sink(A):
    nl = A.dims["nlevsoi"]
    ns = A.dims["nlevsno"]
    q = np.zeros((nl, A.ncol))
    redist = np.zeros(A.npatch)
    for c in range(A.ncol):
        for p in range(A["patchi"][c] - 1, A["patchf"][c]):
            if not (A["active"][p] and A["frac_veg_nosno"][p] > 0 and A["wtcol"][p] > 0):
                continue
            flux = A["k_soil_root"][:nl, p] * (A["smp_l"][:nl, c] - A["vegwp"][3, p] - 1000.0 * A["z"][ns:ns + nl, c])
            redist[p] = flux[flux < 0].sum()
            q[:, c] += flux * A["wtcol"][p]
    return q, redist


def test_root_sink_against_numpy():
    A = PH.synthetic(300, seed=4)
    assert O.phs_root_sink(A) == 0
    q, redist = _numpy_sink(A)
    assert np.allclose(A["qflx_rootsoi"][:20], q, rtol=1e-12, atol=1e-30)
    assert np.allclose(A["qflx_hydr_redist"], redist, rtol=1e-12, atol=1e-30)
    neg = np.where(q < 0, q, 0.0).sum(0)
    assert np.allclose(A["qflx_phs_neg"], neg, rtol=1e-12, atol=1e-30)
    assert np.all(A["qflx_phs_neg"] <= 0)
    excluded = (A["active"] == 0) | (A["frac_veg_nosno"] == 0) | (A["wtcol"] <= 0)
    assert excluded.any() and np.all(A["qflx_hydr_redist"][excluded] == 0.0)


def test_root_sink_vanishes_at_hydrostatic_equilibrium():
    A = PH.synthetic(64, seed=5)
    ns = A.dims["nlevsno"]
    A["z"][ns:] = np.round(A["z"][ns:] * 8) / 8                 # exact in binary64
    A["vegwp"][3] = -np.round(np.abs(A["vegwp"][3]))
    for c in range(A.ncol):
        ps = list(range(A["patchi"][c] - 1, A["patchf"][c]))
        psi = A["vegwp"][3, ps[0]] if ps else -5000.0
        A["vegwp"][3, ps] = psi                                 # one root potential per column
        A["smp_l"][:, c] = psi + 1000.0 * A["z"][ns:, c]
    assert O.phs_root_sink(A) == 0
    assert np.all(A["qflx_rootsoi"][:20] == 0.0)
    assert np.all(A["qflx_phs_neg"] == 0.0) and np.all(A["qflx_hydr_redist"] == 0.0)


# ---- calcstress (PhotosynthesisMod.F90:4490-4710) -------------------------------
RGAS = 6.02214e26 * 1.38065e-23


def _forward_q(A, p, gs, lai):
    """Transpiration a stomatal conductance allows (the Penman form getqflx
    inverts), numpy."""
    c = A["column"][p] - 1
    cf = A["forc_pbot"][c] / (RGAS * 1e-3 * A["thm"][p]) * 1e6
    efpot = A["forc_rho"][c] * (A["elai"][p] + A["esai"][p]) * A["gb_mol"][p] * (A["qsatl"][p] - A["qaf"][p])
    if not (efpot > 0 and A["elai"][p] > 0 and gs > 0):
        return 0.0
    gb = A["gb_mol"][p]
    return efpot * A["fdry"][p] * lai * gs / (gb * gs + gb * gb) * gb / A["elai"][p] / cf


def _plc_np(x, psi50, ck):
    v = 2.0 ** -((x / psi50) ** ck)
    return np.where(v < 0.005, 0.0, v)


def _stress_case(seed=1, ncol=1500):
    A = PH.synthetic_stress(ncol, seed=seed)
    prm = PH.default_params()
    B = A.copy()
    assert O.phs_calcstress(B, prm) == 0
    return A, B, prm


def _seg(prm, ivt):
    p50, ck, km = (k.reshape(4, -1)[:, ivt] for k in prm._keep[4:7])
    return p50, ck, km


def test_calcstress_without_transpiration_is_getvegwp_closed_form():
    A, B, prm = _stress_case(seed=2)
    ns = A.dims["nlevsno"]
    night = A["vegwp"][0] > 0
    dry = (A["qsatl"] <= A["qaf"]) & ~night                     # efpot <= 0: no transpiration
    assert dry.sum() > 20
    for p in np.nonzero(dry)[0]:
        c = A["column"][p] - 1
        k = A["k_soil_root"][:20, p]
        g = 1000.0 * A["z"][ns:ns + 20, c]
        smp = A["smp_l"][:20, c]
        x4 = (k * (smp - g)).sum() / k.sum() if k.sum() != 0 else (smp - g).mean()
        x3 = x4 - 1000.0 * A["htop"][p]
        np.testing.assert_allclose(B["vegwp"][:, p], [x3, x3, x3, x4], rtol=1e-12)
        p50, ck, _ = _seg(prm, A["itype"][p])
        for b, seg in ((B["bsun"][p], 0), (B["bsha"][p], 1)):
            e = float(_plc_np(x3, p50[seg], ck[seg]))
            assert b == pytest.approx(0.0 if e < 0.01 else e, rel=1e-12, abs=0)


def test_calcstress_newton_balances_uptake_and_transpiration():
    A, B, prm = _stress_case(seed=3)
    ns = A.dims["nlevsno"]
    night = A["vegwp"][0] > 0
    checked = candidates = 0
    for p in np.nonzero(~night & (A["qsatl"] > A["qaf"]) & (A["elai"] > 0.01))[0]:
        qs = _forward_q(A, p, A["gs_mol_sun"][p], A["laisun"][p])
        qh = _forward_q(A, p, A["gs_mol_sha"][p], A["laisha"][p])
        if qs + qh <= 0:
            continue
        candidates += 1
        x = B["vegwp"][:, p]
        f, _, _ = O.phs_spac(B, prm, p + 1, x, qs, qh)
        if np.sqrt((f * f).sum()) >= 1e-6 * (qs + qh):
            continue                                            # left by the step or iteration limit
        c = A["column"][p] - 1
        k = A["k_soil_root"][:20, p]
        uptake = (k * (A["smp_l"][:20, c] - x[3] - 1000.0 * A["z"][ns:ns + 20, c])).sum()
        p50, ck, _ = _seg(prm, A["itype"][p])
        lost = qs * _plc_np(x[0], p50[0], ck[0]) + qh * _plc_np(x[1], p50[1], ck[1])
        assert uptake == pytest.approx(lost, rel=1e-5, abs=1e-6 * (qs + qh))
        # the stress factors are the conductances that carry the attenuated fluxes
        if A["laisun"][p] > 0.001 and A["laisha"][p] > 0.001:
            for b, gs, lai, q, seg in ((B["bsun"][p], A["gs_mol_sun"][p], A["laisun"][p], qs, 0),
                                       (B["bsha"][p], A["gs_mol_sha"][p], A["laisha"][p], qh, 1)):
                if q > 0 and b > 0:
                    want = q * _plc_np(x[seg], p50[seg], ck[seg])
                    assert _forward_q(A, p, b * gs, lai) == pytest.approx(want, rel=1e-9)
        checked += 1
    assert checked > 100
    assert checked >= 0.95 * candidates                         # Newton meets tolf = 1e-6 nearly everywhere


@pytest.mark.parametrize("shaded", [True, False])
def test_spacA_inverts_the_jacobian_of_spacF(shaded):
    A, _, prm = _stress_case(seed=4, ncol=300)
    night = A["vegwp"][0] > 0
    ok = ~night & (A["laisun"] > 0.01) & (A["tsai"] > 0.1) & (A["k_soil_root"][:20].sum(0) > 0)
    ok &= (A["laisha"] > 0.01) if shaded else True
    done = 0
    for p in np.nonzero(ok)[0][:40]:
        if not shaded:
            A["laisha"][p] = 0.0
        p50, _, _ = _seg(prm, A["itype"][p])
        x = 0.3 * p50 * np.array([1.3, 1.2, 1.0, 0.8])           # on the smooth part of every curve
        qs, qh = 2e-5, (1e-5 if shaded else 0.0)
        _, iA, flag = O.phs_spac(A, prm, p + 1, x, qs, qh)
        assert flag == 0
        J = np.zeros((4, 4))
        for k in range(4):
            h = 1e-5 * abs(x[k])
            e = np.zeros(4)
            e[k] = h
            fp, _, _ = O.phs_spac(A, prm, p + 1, x + e, qs, qh)
            fm, _, _ = O.phs_spac(A, prm, p + 1, x - e, qs, qh)
            J[:, k] = (fp - fm) / (2 * h)
        # the Newton step is x + invA f, so invA is minus the inverse Jacobian
        if shaded:
            np.testing.assert_allclose(iA @ J, -np.eye(4), atol=1e-5)
        else:
            # spacF swaps f(sun) and f(sha) when laisha is below the limit (the
            # sun equation in the sha slot), and spacA inverts the 3x3 block of
            # slots sha, xyl, root: equations sun, xyl, root against x sun, xyl, root
            Jb = J[np.ix_([1, 2, 3], [0, 2, 3])]
            np.testing.assert_allclose(iA[1:, 1:] @ Jb, -np.eye(3), atol=1e-5)
        done += 1
    assert done >= 10


def test_calcstress_night_outputs_and_ranges():
    A, B, prm = _stress_case(seed=5)
    night = A["vegwp"][0] > 0
    assert night.any() and (~night).any()
    assert np.all(B["qflx_tran_veg"][night] >= 0) and np.all(B["qflx_tran_veg"][~night] == -1.0)
    pd = night & (A["local_morning"] != 0)
    np.testing.assert_array_equal(B["vegwp_pd"][:, pd], B["vegwp"][:, pd])
    assert np.all(B["vegwp_pd"][:, ~pd] == 1e36)
    for b in (B["bsun"], B["bsha"]):
        assert np.all((b == 0) | ((b >= 0.01) & (b <= 1.0 + 1e-9)))


def test_calcstress_night_water_potentials_carry_the_flux():
    # at night getvegwp sets the potentials from the stressed conductances: the
    # soil uptake equals their transpiration, each segment's potential drop
    # carries it through that segment's conductance, and without stem area
    # the xylem sits one canopy height above the root
    A = PH.synthetic_stress(1500, seed=6)
    night = A["vegwp"][0] > 0
    nostem = night & (np.arange(A.npatch) % 5 == 0)
    A["tsai"][nostem] = 0.0
    prm = PH.default_params()
    B = A.copy()
    assert O.phs_calcstress(B, prm) == 0
    checked = 0
    for p in np.nonzero(night)[0]:
        x = B["vegwp"][:, p]
        qs = _forward_q(A, p, B["bsun"][p] * A["gs_mol_sun"][p], A["laisun"][p])
        qh = _forward_q(A, p, B["bsha"][p] * A["gs_mol_sha"][p], A["laisha"][p])
        q = qs + qh
        if A["k_soil_root"][:20, p].sum() > 0:                 # without roots nothing is taken up
            assert B["qflx_tran_veg"][p] == pytest.approx(max(q, 0.0), rel=1e-8, abs=1e-18)
        else:
            assert B["qflx_tran_veg"][p] == 0.0
        p50, ck, km = _seg(prm, A["itype"][p])
        grav1 = 1000.0 * A["htop"][p]
        fr = float(_plc_np(x[3], p50[3], ck[3]))
        if A["tsai"][p] > 0 and fr > 0:
            assert (x[3] - x[2] - grav1) * fr * km[3] / A["htop"][p] * A["tsai"][p] == pytest.approx(q, rel=1e-8,
                                                                                                   abs=1e-18)
        else:
            assert x[2] == pytest.approx(x[3] - grav1, rel=1e-14)
        fx = float(_plc_np(x[2], p50[2], ck[2]))
        if A["laisha"][p] > 0 and fx > 0:
            assert (x[2] - x[1]) * fx * km[2] * A["laisha"][p] == pytest.approx(qh, rel=1e-8, abs=1e-18)
        checked += 1
    assert checked > 300 and nostem.any()
