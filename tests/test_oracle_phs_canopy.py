"""Pins of the oracle's PhotosynthesisHydraulicStress (the coupled sun / shade
ci solve with calcstress inside, PhotosynthesisMod.F90:2704-4486) and of
CanopyFluxes with plant hydraulic stress -- CPU only.

The oracle's restatement is checked against what the mathematics fixes, each
evaluated here independently in numpy from the call's outputs:
  - the transpiration the hydraulics deliver equals the leaves' water loss at
    the solved stomatal conductances (getqflx's Penman form,
    :5080-5164), and the root water potential is the one at which the soil
    supplies exactly that (getvegwp, :4979-5077): sum_j k_j (smp_j - psi_root -
    1000 z_j) = transpiration;
  - the sun / shade flux balance of the leaf segments at the final potentials:
    q_leaf = plc(psi_xyl) kmax laisun (psi_xyl - psi_leaf);
  - the Ball-Berry identity at the solution of the coupled ci solve with the
    stress factor: gs = m an hs / cs p + max(b bbb, 1) (the check the
    reference logs at :3677-3695);
  - the night branch: no photosynthesis, stress factors from the minimum
    conductance, the predawn potential written only in the morning;
  - without transpiration the stress factors are the vulnerability curve at
    the final potentials (the curve whose known answer
    test_Photosynthesis.pf:65-70 gives), the leaves at the xylem's potential;
  - CanopyFluxes with PHS: the leaf energy residual within 0.1 W/m2, btran
    the leaf-area weighted bsun / bsha, transpiration not above what the
    hydraulics deliver.
"""
import numpy as np
import pytest

from ctsm_amd import phs as PH, synthetic as S
from oracle import oracle as O

RGAS = 6.02214e26 * 1.38065e-23


def _case(seed, stomatal, night_frac=0.0, ncol=96):
    st = S.canopy_workload("canopy", ncol=ncol)
    A = PH.canopy_phs(st, seed, night_frac=night_frac)
    npft = len(S.pft_table()["dleaf"])
    pp = PH.default_params(npft)
    cfg = O.canopy_config(stomatalcond_method=stomatal)
    assert O.phs_conductance_for(st, A, pp, cfg) == 0
    return st, A, pp, cfg


def _plc(x, psi50, ck):
    v = 2.0 ** (-(x / psi50) ** ck)
    return np.where(v < 0.005, 0.0, v)


def _table(pp, name, npft):
    return np.ctypeslib.as_array(getattr(pp, name), shape=(4 * npft,)).reshape(4, npft)


def _run_patches(st, A, pp, cfg, rng, patches):
    """Call the oracle's PHS photosynthesis for each patch at canopy-air
    values the iteration could pass; returns the inputs used."""
    prm, pft = S.canopy_params(), O.PftStruct(S.pft_table())
    used = {}
    for p in patches:
        i = p - 1
        c = st.pch["column"][i] - 1
        pbot = st.can["forc_pbot"][c]
        qs, es, _ = O.qsat(np.array([st.can["t_veg"][i]]), pbot)
        qsatl, esat = float(qs[0]), float(es[0])
        qaf = qsatl * rng.uniform(0.4, 0.95)
        eair = pbot * qaf / 0.622
        rb = rng.uniform(15.0, 80.0)
        rc = O.photosynthesis_phs_patch(st, A, cfg, prm, pft, pp, p, esat, eair, 0.209 * pbot, 4.0e-4 * pbot, rb,
                                        1.0, qsatl, qaf)
        assert rc == 0, p
        used[p] = dict(qsatl=qsatl, qaf=qaf, esat=esat, eair=eair, rb=rb, pbot=pbot, c=c)
    return used


def _getqflx(st, i, c, gb, gs, lai, qsatl, qaf, thm):
    cf = st.can["forc_pbot"][c] / (RGAS * 1.e-3 * thm) * 1.e6
    elai = st.can["elai"][i]
    efpot = st.can["forc_rho"][c] * (elai + st.can["esai"][i]) * gb * (qsatl - qaf)
    if not (efpot > 0 and elai > 0 and gs > 0):
        return 0.0
    return efpot * (st.can["fdry"][i] / gb * (lai / (1.0 / gb + 1.0 / gs)) / elai) / cf


@pytest.mark.parametrize("stomatal", [2, 1], ids=["medlyn", "ballberry"])
def test_transpiration_equals_root_supply_and_leaf_segments_balance(stomatal):
    st, A, pp, cfg = _case(5, stomatal)
    rng = np.random.default_rng(9)
    lit = np.nonzero((st.can["nrad"] >= 1) & (st.can["parsun_z"] > 0) & (st.can["elai"] > 0.5))[0][:60] + 1
    used = _run_patches(st, A, pp, cfg, rng, lit)
    npft = len(S.pft_table()["dleaf"])
    kmax, psi50, ck = (_table(pp, k, npft) for k in ("kmax", "psi50", "ck"))
    ns, nsoi = st.dims.nlevsno, st.dims.nlevsoi
    checked = 0
    for p, u in used.items():
        i, c = p - 1, u["c"]
        gb = st.can["gb_mol"][i]
        gs_sun, gs_sha = st.can["gs_mol_sun"][i], st.can["gs_mol_sha"][i]
        if A["an_sun"][i] < 0 or A["an_sha"][i] < 0:
            continue          # the conductance then is the minimum one, set after the solve
        thm = st.can["thm"][i]
        q_sun = _getqflx(st, i, c, gb, gs_sun, st.can["laisun"][i], u["qsatl"], u["qaf"], thm)
        q_sha = _getqflx(st, i, c, gb, gs_sha, st.can["laisha"][i], u["qsatl"], u["qaf"], thm)
        tran = st.can["qflx_tran_veg"][i]
        k = A["k_soil_root"][:nsoi, i]
        if k.sum() <= 0:
            continue
        smp = st.col["smp_l"][:nsoi, c]
        z = st.col["z"][ns:ns + nsoi, c]
        supply = np.sum(k * (smp - A["vegwp"][3, i] - 1000.0 * z))
        assert tran == pytest.approx(max(q_sun + q_sha, 0.0), rel=1e-9, abs=1e-18), p
        assert supply == pytest.approx(tran, rel=1e-9, abs=1e-15), p
        # the leaf segments carry their fluxes from the xylem (getvegwp)
        ivt = st.can["itype"][i]
        xyl = A["vegwp"][2, i]
        fx = _plc(xyl, psi50[2, ivt], ck[2, ivt])
        for q, lai, seg in ((q_sun, st.can["laisun"][i], 0), (q_sha, st.can["laisha"][i], 1)):
            if lai > 0 and fx > 0:
                assert q == pytest.approx(fx * kmax[2, ivt] * lai * (xyl - A["vegwp"][seg, i]), rel=1e-9), (p, seg)
        checked += 1
    assert checked >= 20


def test_ball_berry_identity_at_the_coupled_solution():
    st, A, pp, cfg = _case(6, 1)
    rng = np.random.default_rng(10)
    lit = np.nonzero((st.can["nrad"] >= 1) & (st.can["parsun_z"] > 0) & (st.can["elai"] > 0.5))[0][:80] + 1
    used = _run_patches(st, A, pp, cfg, rng, lit)
    mbb = st.can["mbb"]
    bbb = st.can["bbb"]
    total = stressed = 0
    for p, u in used.items():
        i = p - 1
        for ph, an, b in (("sun", A["an_sun"][i], A["bsun"][i]), ("sha", A["an_sha"][i], A["bsha"][i])):
            if an <= 0 or b <= 0:
                continue
            gs = st.can["gs_mol_" + ph][i]
            gb = st.can["gb_mol"][i]
            ceair = min(u["eair"], u["esat"])
            hs = (gb * ceair + gs * u["esat"]) / ((gb + gs) * u["esat"])
            cs = max(4.0e-4 * u["pbot"] - 1.4 / gb * an * u["pbot"], 1.e-6)
            gs_bb = mbb[i] * an * hs / cs * u["pbot"] + max(b * bbb[i], 1.0)
            # gs is the root of the quadratic at the final an, and hs, cs are
            # functions of (an, gs): the identity holds to rounding wherever
            # the solve stopped, converged or not
            assert gs == pytest.approx(gs_bb, rel=1e-12), (p, ph)
            total += 1
            stressed += b < 0.9
    # stressed leaves tell max(b bbb, 1) from max(bbb, 1) by more than 1 %
    assert total >= 30 and stressed >= 3, (total, stressed)


def test_night_branch():
    st, A, pp, cfg = _case(7, 2, night_frac=1.0)
    rng = np.random.default_rng(11)
    dark = np.nonzero((st.can["nrad"] >= 1) & (st.can["parsun_z"] <= 0))[0][:40] + 1
    assert len(dark) >= 20
    used = _run_patches(st, A, pp, cfg, rng, dark)
    npft = len(S.pft_table()["dleaf"])
    psi50, ck = _table(pp, "psi50", npft), _table(pp, "ck", npft)
    for p in used:
        i = p - 1
        ivt = st.can["itype"][i]
        assert np.all(A["ac_phs"][:, i] == 0.0) and np.all(A["ag_phs"][:, i] == 0.0)
        assert st.can["psnsun"][i] == 0.0 and st.can["psnsha"][i] == 0.0
        # an = ag - b lmr with no photosynthesis: the respiration's sign
        assert A["an_sun"][i] == -A["bsun"][i] * st.can["lmrsun_z"][i]
        # predawn potential only in the morning
        if A["local_morning"][i]:
            np.testing.assert_array_equal(A["vegwp_pd"][:, i], A["vegwp"][:, i])
        else:
            assert np.all(A["vegwp_pd"][:, i] == 1.e36)
        # where Newton gave up (or nothing transpires) the factors are plc of the potentials
        for seg, b in ((0, A["bsun"][i]), (1, A["bsha"][i])):
            pl = float(_plc(A["vegwp"][seg, i], psi50[seg, ivt], ck[seg, ivt]))
            assert b == 0.0 or 0.0 < b <= 1.0
            if st.can["laisun"][i] <= 0.001 and st.can["laisha"][i] <= 0.001:
                assert b == (pl if pl >= 0.01 else 0.0)


def test_stress_factor_is_the_curve_at_the_final_potential_without_transpiration():
    """With no leaf area nothing transpires: calcstress skips Newton, getvegwp
    sets the potentials and bsun = plc(psi_sun) on the reference's curve
    (psi50 = -150000 mm, ck = 3.95, whose value at -25000 mm is
    test_Photosynthesis.pf's known answer 0.999415208562283)."""
    st, A, pp, cfg = _case(8, 2, night_frac=1.0, ncol=16)
    npft = len(S.pft_table()["dleaf"])
    psi50 = _table(pp, "psi50", npft)
    ck = _table(pp, "ck", npft)
    psi50[:] = -150000.0
    ck[:] = 3.95
    assert abs(O.phs_plc(-25000.0, -150000.0, 3.95) - 0.999415208562283) <= 1e-13
    idx = np.nonzero(st.can["nrad"] >= 1)[0][:6]
    for i in idx:
        st.can["laisun"][i] = st.can["laisha"][i] = 0.0
    rng = np.random.default_rng(1)
    _run_patches(st, A, pp, cfg, rng, idx + 1)
    for i in idx:
        for seg, b in ((0, A["bsun"][i]), (1, A["bsha"][i])):
            expect = float(_plc(A["vegwp"][seg, i], -150000.0, 3.95))
            assert b == (expect if expect >= 0.01 else 0.0), (i, seg)
        # no leaves: the leaf potentials are the xylem's
        assert A["vegwp"][0, i] == A["vegwp"][2, i] and A["vegwp"][1, i] == A["vegwp"][2, i]


@pytest.mark.parametrize("stomatal", [2, 1], ids=["medlyn", "ballberry"])
def test_canopyfluxes_with_phs_energy_residual_and_btran(stomatal):
    st = S.canopy_workload("canopy_uneven", ncol=128)
    A = PH.canopy_phs(st, 3, night_frac=0.3)
    npft = len(S.pft_table()["dleaf"])
    pp = PH.default_params(npft)
    cfg = O.canopy_config(stomatalcond_method=stomatal)
    rc, bad = O.canopyfluxes_phs(st, A, cfg, S.canopy_params(), O.PftStruct(S.pft_table()), pp)
    assert rc == 0, bad
    assert np.max(np.abs(st.can["canopy_err"])) <= 0.1
    lsun, lsha = st.can["laisun_z"], st.can["laisha_z"]
    nr = st.can["nrad"] >= 1
    w = nr & ((lsun + lsha) > 0)
    bt = A["bsun"] * (lsun / (lsun + np.where(w, lsha, 1.0))) + A["bsha"] * (lsha / (lsun + np.where(w, lsha, 1.0)))
    np.testing.assert_allclose(st.can["btran"][w], bt[w], rtol=1e-15, atol=0)
    assert np.all((A["bsun"] >= 0) & (A["bsun"] <= 1)) and np.all((A["bsha"] >= 0) & (A["bsha"] <= 1))
    assert np.all(st.can["qflx_tran_veg"] >= 0)
    # some patches are stressed, some are not
    assert np.any(A["bsun"][nr] < 0.9) and np.any(A["bsun"][nr] > 0.95)
