"""Pins of the oracle functions that bit-equality with the GPU cannot fix:
closed forms and mathematical properties that a misread formula fails.

Each check here is independent of the oracle's own code path:
  * SoilThermProp (SoilTemperatureMod.F90:602-901): the layer conductivity
    `thk` at the saturated unfrozen / saturated frozen / dry limits of the
    Johansen-Farouki model (k_sat = k_solids^(1-n) k_liq^n or k_ice^n), the
    Kersten number's two branches and its switch at tfrz, bedrock; the heat
    capacity through fact = dtime / cv (and the capr correction of the top
    layer).
  * The discrete heat equation: a steady conduction profile through a layered
    column (T_{j+1} - T_j = F [(zi_j - z_j)/k_j + (z_{j+1} - zi_j)/k_{j+1}])
    with the matching surface and bottom fluxes is a fixed point of the
    Crank-Nicolson step -- this fixes the interface conductance (series
    resistances), the flux signs and every layer's k.
  * Phasechange's supercooled water (SoilTemperatureMod.F90:1323-1335,
    1451-1456): a frozen layer holding more liquid than the freezing-point
    depression allows keeps exactly watsat (smp/sucsat)^(-1/b) dz 1000 with
    smp = hfus (tfrz - T) / (g T) 1000.
  * Photosynthesis kinetics (PhotosynthesisMod.F90:1525-1846, 2488-2561):
    ft is an Arrhenius ratio equal to 1 at 25 C, fth is normalised to 1 at 25
    C, Kc / Ko / Gamma* equal their 25 C values there, Vcmax(25 C) equals
    N_a F_LNR F_NR a_R25 times the canopy scaling, each rate follows its own
    activation / deactivation energy with temperature, the C4 Q10 form.
  * The co-limitation (ci_func, :2564-2701) reduces to min(Wc, Wj, Wp) when
    both curvatures are 1 and stays below it otherwise.
  * Roughness and displacement (CanopyFluxesMod.F90:897-943): the Zeng-Wang
    limits (bare soil at lt = 0, unchanged at lt >= tlsai_crit) and the
    Raupach (1994) / Meier (2022) forms with U/u* the root of x = x0 e^x.
  * Resistances (CanopyFluxesMod.F90:1016-1047): u_af = u*,
    the leaf boundary layer r_b = (d_leaf / u_af)^(1/2) / C_v, the under-canopy
    soil resistance 1 / (C_s u_af) with the Sakaguchi-Zeng C_s.
  * QSat (QSatMod.F90:61-127): the Flatau et al. (1992) 8th-order fits
    evaluated in exact rational arithmetic from the published coefficients.

tests/test_oracle_mutants.py plants mistakes in copies of the oracle and
checks that these pins (with the older ones) catch each of them.
"""
from fractions import Fraction

import numpy as np
import pytest

from ctsm_amd import synthetic as S
from oracle import oracle as O

# physical constants (shr_const_mod / clm_varcon published values)
TFRZ, HFUS, GRAV, SB = 273.15, 3.337e5, 9.80616, 5.67e-8
TKWAT, TKICE, TKBED, CSOL_BED = 0.57, 2.29, 3.0, 2.0e6
CPLIQ, CPICE, CAPR = 4.188e3, 2.11727e3, 0.34
RGAS = 6.02214e26 * 1.38065e-23 * 1e-3      # J/K/mol
T25 = TFRZ + 25.0
DT = 1800.0


def _soil_state(n, seed=0xC75D00A0, nbedrock=None):
    """Soil-only columns (no snow, no standing water, one bare patch) whose
    surface forcing is set by the tests through sabg_lyr of level 1."""
    st = S.soil_only(n, seed)
    c, p = st.col, st.pch
    for k in ("frac_sno_eff", "frac_h2osfc", "h2osfc", "h2osno_no_layers", "forc_lwrad", "eflx_bot"):
        c[k][:] = 0.0
    c["emg"][:] = 0.97
    if nbedrock is not None:
        c["nbedrock"][:] = nbedrock
    for k in ("sabg", "sabg_soil", "sabg_snow", "dlrad", "cgrnd", "eflx_sh_grnd", "eflx_sh_snow",
              "eflx_sh_soil", "eflx_sh_h2osfc", "qflx_evap_soi", "qflx_ev_snow", "qflx_ev_soil",
              "qflx_ev_h2osfc"):
        p[k][:] = 0.0
    p["frac_veg_nosno"][:] = 0
    p["wtcol"][:] = 1.0
    p["sabg_lyr"][...] = 0.0
    return st


def _soil(st, name):
    """(nlevgrnd, ncol) view of a [SNO] field's soil rows."""
    return st.col[name][st.dims.nlevsno:]


def _set_surface_flux(st, hs):
    """Surface heat flux into the top soil layer of snow-free columns: with no
    longwave in, no turbulent fluxes and frac_veg_nosno = 0 the top layer gets
    sabg_lyr(1) - emg sigma T1^4."""
    T1 = _soil(st, "t_soisno")[0]
    st.pch["sabg_lyr"][st.dims.nlevsno] = hs + st.col["emg"] * SB * T1 ** 4
    st.col["t_grnd"][:] = T1
    st.col["t_h2osfc"][:] = T1


def _thk_closed(watsat, tkmg, tkdry, satw, fl, unfrozen):
    """Johansen / Farouki: k = Ke k_sat + (1 - Ke) k_dry with
    k_sat = tkmg k_liq^(fl n) k_ice^((1-fl) n); Ke = max(0, log10 S + 1)
    unfrozen, S frozen; k_dry below S = 1e-7."""
    ksat = tkmg * np.power(TKWAT, fl * watsat) * np.power(TKICE, (1.0 - fl) * watsat)
    ke = np.where(unfrozen, np.maximum(0.0, np.log10(np.maximum(satw, 1e-300)) + 1.0), satw)
    return np.where(satw > 1e-7, ke * ksat + (1.0 - ke) * tkdry, tkdry)


# ---------------------------------------------------------------------------
# SoilThermProp

REGIMES = [  # (saturation, liquid fraction of the water, temperature K)
    (1.0, 1.0, 290.0),      # saturated unfrozen: tkmg * 0.57^n
    (1.0, 0.0, 260.0),      # saturated frozen:   tkmg * 2.29^n
    (0.0, 1.0, 290.0),      # dry: tkdry
    (0.5, 1.0, 290.0),      # Kersten log branch
    (0.5, 0.0, 265.0),      # Kersten frozen branch: Ke = S
    (0.05, 1.0, 290.0),     # log10 S + 1 < 0 clamped: tkdry
    (0.6, 0.5, 272.0),      # mixed liquid / ice below tfrz
    (0.5, 1.0, TFRZ),       # exactly at tfrz: the unfrozen branch
]


def _regime_state(nbedrock=None):
    n = len(REGIMES)
    st = _soil_state(n, nbedrock=nbedrock)
    NG = st.dims.nlevgrnd
    dz = _soil(st, "dz")
    ws = st.col["watsat"]
    liq, ice, t = _soil(st, "h2osoi_liq"), _soil(st, "h2osoi_ice"), _soil(st, "t_soisno")
    for c, (sat, fl, T) in enumerate(REGIMES):
        vol = sat * ws[:, c]
        liq[:, c] = fl * vol * dz[:, c] * 1000.0
        ice[:, c] = (1.0 - fl) * vol * dz[:, c] * 917.0
        t[:, c] = T
    _set_surface_flux(st, 0.0)
    return st, NG


@pytest.mark.parametrize("nbedrock", [None, 9])
def test_thk_closed_forms(nbedrock):
    st, NG = _regime_state(nbedrock)
    liq0, ice0, t0 = (_soil(st, k).copy() for k in ("h2osoi_liq", "h2osoi_ice", "t_soisno"))
    dz, ws = _soil(st, "dz"), st.col["watsat"]
    rc, _ = O.soiltemperature(st)
    assert rc == 0
    thk = _soil(st, "thk")
    vliq, vice = liq0 / (1000.0 * dz), ice0 / (917.0 * dz)
    satw = np.minimum(1.0, (vliq + vice) / ws)
    with np.errstate(invalid="ignore", divide="ignore"):
        fl = np.where(vliq + vice > 0, vliq / (vliq + vice), 1.0)
    want = _thk_closed(ws, st.col["tkmg"], st.col["tkdry"], satw, fl, t0 >= TFRZ)
    nb = st.col["nbedrock"]
    j = np.arange(1, NG + 1)[:, None]
    want = np.where(j > nb[None, :], TKBED, want)
    np.testing.assert_allclose(thk, want, rtol=1e-13, atol=0)
    # the named limits, spelled out
    c_sat, c_frz, c_dry = 0, 1, 2
    soil = j[:, 0] <= nb[0]
    np.testing.assert_allclose(thk[soil, c_sat], (st.col["tkmg"] * TKWAT ** ws)[soil, c_sat], rtol=1e-13)
    np.testing.assert_allclose(thk[soil, c_frz], (st.col["tkmg"] * TKICE ** ws)[soil, c_frz], rtol=1e-13)
    np.testing.assert_array_equal(thk[soil, c_dry], st.col["tkdry"][soil, c_dry])
    assert np.all(thk[~soil] == TKBED)


@pytest.mark.parametrize("nbedrock", [None, 9])
def test_heat_capacity_through_fact(nbedrock):
    st, NG = _regime_state(nbedrock)
    liq0, ice0 = _soil(st, "h2osoi_liq").copy(), _soil(st, "h2osoi_ice").copy()
    dz, z = _soil(st, "dz"), _soil(st, "z")
    rc, _ = O.soiltemperature(st)
    assert rc == 0
    fact = _soil(st, "fact")
    cv = st.col["csol"] * (1.0 - st.col["watsat"]) * dz + CPICE * ice0 + CPLIQ * liq0
    j = np.arange(1, NG + 1)[:, None]
    cv = np.where(j > st.col["nbedrock"][None, :], CSOL_BED * dz, cv)
    np.testing.assert_allclose(fact[1:], DT / cv[1:], rtol=1e-13)
    # top layer: dt / cv * dz / (0.5 (z1 - zi0 + capr (z2 - zi0))), zi0 = 0
    top = DT / cv[0] * dz[0] / (0.5 * (z[0] + CAPR * z[1]))
    np.testing.assert_allclose(fact[0], top, rtol=1e-13)


def test_snow_heat_capacity_through_fact():
    st = S.soil_only(64, 0xC75D00A1, snow=True)
    rc, _ = O.soiltemperature(st.copy())
    c = st.col
    NS = st.dims.nlevsno
    o = st.copy()
    rc, _ = O.soiltemperature(o)
    assert rc == 0
    snl = c["snl"]
    checked = 0
    for col in range(st.dims.ncol):
        for j in range(snl[col] + 2, 1):          # snow layers below the top one
            r = NS - 1 + j
            cv = max(1e-6, (CPLIQ * c["h2osoi_liq"][r, col] + CPICE * c["h2osoi_ice"][r, col]) /
                     c["frac_sno_eff"][col])
            assert o.col["fact"][r, col] == pytest.approx(DT / cv, rel=1e-13)
            checked += 1
    assert checked > 100


@pytest.mark.parametrize("flux", [-2.0, -0.7])
def test_steady_conduction_profile_is_fixed_point(flux):
    """Constant upward heat flux through layers of different conductivity
    (saturated and half-saturated soil, bedrock below a varying depth): the
    node temperatures of the exact series-resistance profile stay put."""
    n = 24
    st = _soil_state(n, seed=0xC75D00A2)
    rng = np.random.default_rng(7)
    st.col["nbedrock"][:] = rng.integers(5, 21, n)
    NG = st.dims.nlevgrnd
    dz, z = _soil(st, "dz"), _soil(st, "z")
    zi = st.col["zi"][st.dims.nlevsno:]            # zi(0 .. nlevgrnd) rows
    ws = st.col["watsat"]
    sat = np.where(rng.random((NG, n)) < 0.5, 1.0, 0.5)
    liq = _soil(st, "h2osoi_liq")
    liq[...] = sat * ws * dz * 1000.0
    _soil(st, "h2osoi_ice")[...] = 0.0
    k = _thk_closed(ws, st.col["tkmg"], st.col["tkdry"], sat, np.ones_like(sat), np.ones(sat.shape, bool))
    k = np.where(np.arange(1, NG + 1)[:, None] > st.col["nbedrock"][None, :], TKBED, k)
    # series resistance between nodes j and j+1 (zi row j is the interface below layer j)
    R = (zi[1:NG] - z[:NG - 1]) / k[:NG - 1] + (z[1:] - zi[1:NG]) / k[1:]
    T = np.empty((NG, n))
    T[0] = 280.0
    T[1:] = 280.0 - flux * np.cumsum(R, axis=0)    # upward flux (flux < 0): warmer below
    _soil(st, "t_soisno")[...] = T
    assert T.min() > TFRZ + 1.0
    _set_surface_flux(st, flux)                  # heat into the top = flux
    st.col["eflx_bot"][:] = -flux                # heat into the bottom layer from below
    rc, _ = O.soiltemperature(st)
    assert rc == 0
    np.testing.assert_allclose(_soil(st, "t_soisno"), T, rtol=0, atol=2e-9)


def test_supercooled_water_limit():
    """Isothermal frozen columns with no heat flux: every layer holding more
    liquid than the freezing-point depression allows and cold enough to stay
    frozen after refreezing keeps exactly the supercooled amount."""
    n = 16
    st = _soil_state(n, seed=0xC75D00A3)
    T0 = np.linspace(255.0, 270.0, n)
    dz = _soil(st, "dz")
    ws, bsw, sucsat = st.col["watsat"], st.col["bsw"], st.col["sucsat"]
    smp = HFUS * (TFRZ - T0) / (GRAV * T0) * 1000.0
    sc = ws * np.power(smp[None, :] / sucsat, -1.0 / bsw) * dz * 1000.0
    assert np.all(sc < ws * dz * 1000.0)
    _soil(st, "t_soisno")[...] = T0[None, :]
    _soil(st, "h2osoi_liq")[...] = sc + 0.02 * dz * 1000.0
    _soil(st, "h2osoi_ice")[...] = 0.4 * ws * dz * 917.0
    _set_surface_flux(st, 0.0)
    rc, _ = O.soiltemperature(st)
    assert rc == 0
    t1 = _soil(st, "t_soisno")
    frozen = t1 < TFRZ - 1e-6
    assert frozen.sum() > 0.8 * frozen.size
    np.testing.assert_allclose(_soil(st, "h2osoi_liq")[frozen], sc[frozen], rtol=1e-11)
    assert np.all(st.col["imelt"][st.dims.nlevsno:][frozen] == 2)


# ---------------------------------------------------------------------------
# Photosynthesis kinetics and co-limitation

def _psn_setup(n=6, **prm_over):
    st = S.canopy_workload("canopy", ncol=n, seed=0xC75D00A4)
    prm = S.canopy_params()
    prm.update(prm_over)
    pft = S.pft_table()
    return st, prm, pft


def _run_psn(st, prm, pft, p, T, sun=1, cfg=None, btran=0.8, dayl_factor=0.9):
    cfg = cfg or O.canopy_config()
    g = st.can
    c = st.pch["column"][p - 1] - 1
    pbot = g["forc_pbot"][c]
    g["t_veg"][p - 1] = T
    qs, es = S._qsat_np(np.array([T]), np.array([pbot]))
    oair, cair = 0.209 * pbot, 367e-6 * pbot
    rc = O.photosynthesis_patch(st, cfg, prm, O.PftStruct(pft), p, sun, float(es[0]), 0.6 * float(es[0]),
                                oair, cair, 20.0, btran, dayl_factor)
    return rc, pbot, oair


def _lit_patches(st, pft, c3=True):
    it = st.can["itype"]
    lit = (st.can["parsun_z"] > 10.0) & (st.can["nrad"] >= 1)
    is_c3 = pft["c3psn"][it] == 1.0
    return [int(i) + 1 for i in np.nonzero(lit & (is_c3 if c3 else ~is_c3))[0]]


def test_temperature_functions():
    for ha in (79430.0, 36380.0, 72000.0, 50000.0):
        assert O.ft(T25, ha) == 1.0
        for T in (260.0, 285.0, 310.0, 320.0):
            # Arrhenius: ln f = Ha / R (1/T25 - 1/T)
            assert O.ft(T, ha) == pytest.approx(np.exp(ha / RGAS * (1.0 / T25 - 1.0 / T)), rel=1e-13)
    for hd, se in ((200000.0, 650.0), (150650.0, 490.0)):
        c = O.fth25(hd, se)
        assert c == pytest.approx(1.0 + np.exp((se * T25 - hd) / (RGAS * T25)), rel=1e-14)
        assert O.fth(T25, hd, se, c) == pytest.approx(1.0, abs=2e-16)
        for T in (300.0, 315.0, 325.0):
            assert O.fth(T, hd, se, c) < 1.0
            assert O.fth(T, hd, se, c) == pytest.approx(c / (1.0 + np.exp((se * T - hd) / (RGAS * T))), rel=1e-13)


def test_kinetic_constants_at_25C_and_Vcmax25():
    st, prm, pft = _psn_setup()
    ps = _lit_patches(st, pft)
    assert len(ps) >= 5
    cfg = O.canopy_config()
    for p in ps[:8]:
        btran, dfac = 0.8, 0.9
        rc, pbot, oair = _run_psn(st, prm, pft, p, T25, btran=btran, dayl_factor=dfac, cfg=cfg)
        assert rc == 0
        g = st.can
        assert g["kc"][p - 1] == pytest.approx(prm["kc25_coef"] * pbot, rel=1e-15)
        assert g["ko"][p - 1] == pytest.approx(prm["ko25_coef"] * pbot, rel=1e-15)
        # Gamma*(25) = 0.5 O / (0.5 * 0.209 / cp25) = cp25 * O / 0.209
        assert g["cp"][p - 1] == pytest.approx(prm["cp25_yr2000"] * oair / 0.209, rel=1e-14)
        it = g["itype"][p - 1]
        na = 1.0 / (pft["slatop"][it] * pft["leafcn"][it])
        v25 = na * pft["flnr"][it] * prm["fnr"] * prm["act25"] * dfac * pft["fnitr"][it]
        want = v25 * g["vcmaxcintsun"][p - 1] * btran
        assert g["vcmax_z"][p - 1] == pytest.approx(want, rel=1e-13)
        # TPU carries no soil-moisture stress (only Vcmax and leaf respiration are scaled by btran)
        assert g["tpu_z"][p - 1] == pytest.approx(prm["tpu25ratio"] * want / btran, rel=1e-13)


@pytest.mark.parametrize("T", [283.0, 305.0, 318.0])
def test_rates_follow_their_own_activation_energies(T):
    st, prm, pft = _psn_setup()
    ps = _lit_patches(st, pft)[:6]
    for p in ps:
        g = st.can
        t10c = min(max(g["t_a10"][p - 1] - TFRZ, 11.0), 35.0)
        out = {}
        for TT in (T25, T):
            rc, _, _ = _run_psn(st, prm, pft, p, TT)
            assert rc == 0
            out[TT] = (g["vcmax_z"][p - 1], g["tpu_z"][p - 1], g["lmrsun_z"][p - 1])

        def ratio(ha, hd, se):
            c = 1.0 + np.exp((se * T25 - hd) / (RGAS * T25))
            return np.exp(ha / RGAS * (1.0 / T25 - 1.0 / T)) * c / (1.0 + np.exp((se * T - hd) / (RGAS * T)))

        se_v = (668.39 - 1.07 * t10c) * prm["vcmaxse_sf"]
        se_t = (668.39 - 1.07 * t10c) * prm["tpuse_sf"]
        assert out[T][0] / out[T25][0] == pytest.approx(ratio(prm["vcmaxha"], prm["vcmaxhd"], se_v), rel=1e-12)
        assert out[T][1] / out[T25][1] == pytest.approx(ratio(prm["tpuha"], prm["tpuhd"], se_t), rel=1e-12)
        assert out[T][2] / out[T25][2] == pytest.approx(ratio(prm["lmrha"], prm["lmrhd"], prm["lmrse"]), rel=1e-12)


def test_c4_vcmax_q10_form():
    st, prm, pft = _psn_setup(n=24)
    ps = _lit_patches(st, pft, c3=False)
    assert ps
    for p in ps[:4]:
        vals = {}
        for T in (T25, 303.0):
            rc, _, _ = _run_psn(st, prm, pft, p, T)
            assert rc == 0
            vals[T] = st.can["vcmax_z"][p - 1]

        def q10(T):
            return 2.0 ** ((T - T25) / 10.0) / (1.0 + np.exp(0.2 * (TFRZ + 15.0 - T))) / \
                (1.0 + np.exp(0.3 * (T - (TFRZ + 40.0))))

        assert vals[303.0] / vals[T25] == pytest.approx(q10(303.0) / q10(T25), rel=1e-12)


def test_colimitation_limits():
    st, prm, pft = _psn_setup(n=12)
    ps = _lit_patches(st, pft)
    # both curvatures 1: the co-limited rate is the smallest of the three
    prm1 = dict(prm, theta_ip=1.0)
    pft1 = dict(pft, theta_cj=np.ones_like(pft["theta_cj"]))
    for p in ps[:10]:
        for T in (285.0, 300.0, 312.0):
            rc, _, _ = _run_psn(st, prm1, pft1, p, T)
            assert rc == 0
            g = st.can
            ac, aj, ap, ag = g["ac"][p - 1], g["aj"][p - 1], g["ap"][p - 1], g["ag"][p - 1]
            assert ag == pytest.approx(min(ac, aj, ap), rel=1e-6)
            assert ap == pytest.approx(3.0 * g["tpu_z"][p - 1], rel=1e-15)
    # only the light/Rubisco curvature at 1: the intermediate rate is min(Wc, Wj)
    # and the co-limited rate the smaller root of theta_ip A^2 - (Wi + Wp) A + Wi Wp
    th = prm["theta_ip"]
    assert th < 1.0
    for p in ps[:10]:
        rc, _, _ = _run_psn(st, prm, pft1, p, 300.0)
        assert rc == 0
        g = st.can
        ac, aj, ap, ag = g["ac"][p - 1], g["aj"][p - 1], g["ap"][p - 1], g["ag"][p - 1]
        wi = min(ac, aj)
        b = wi + ap
        root = (b - np.sqrt(b * b - 4.0 * th * wi * ap)) / (2.0 * th)
        assert ag == pytest.approx(root, rel=1e-6)
    # curvatures below 1: strictly below the minimum, and positive
    for p in ps[:10]:
        rc, _, _ = _run_psn(st, prm, pft, p, 300.0)
        g = st.can
        ac, aj, ap, ag = g["ac"][p - 1], g["aj"][p - 1], g["ap"][p - 1], g["ag"][p - 1]
        if min(ac, aj, ap) > 0:
            assert 0.0 < ag < min(ac, aj, ap)


# ---------------------------------------------------------------------------
# Roughness, displacement and resistances

def _canopy_run(n=32, seed=0xC75D00A5, **cfg_over):
    st = S.canopy_workload("canopy", ncol=n, seed=seed)
    cfg = O.canopy_config(**cfg_over)
    before = st.copy()
    rc, bad = O.canopyfluxes(st, cfg, S.canopy_params(), O.PftStruct(S.pft_table()))
    assert rc == 0, bad
    return before, st


def test_zengwang_roughness_limits():
    st0 = S.canopy_workload("canopy", ncol=32, seed=0xC75D00A5)
    g = st0.can
    npch = st0.dims.npatch
    lt_set = np.array([0.0, 0.3, 1.0, 2.0, 5.0])
    lt = lt_set[np.arange(npch) % len(lt_set)]
    g["elai"][:] = 0.7 * lt
    g["esai"][:] = 0.3 * lt
    before = st0.copy()
    cfg = O.canopy_config(z0param_method=0)
    rc, _ = O.canopyfluxes(st0, cfg, S.canopy_params(), O.PftStruct(S.pft_table()))
    assert rc == 0
    a = st0.can
    z0mg = st0.can["z0mg"][st0.pch["column"] - 1]
    d_in, z_in = before.can["displa"], before.can["z0mv"]
    ltc = np.minimum(g["elai"] + g["esai"], 2.0)
    egvf = (1.0 - np.exp(-ltc)) / (1.0 - np.exp(-2.0))
    np.testing.assert_allclose(a["displa"], egvf * d_in, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(a["z0mv"], np.exp(egvf * np.log(z_in) + (1 - egvf) * np.log(z0mg)), rtol=1e-12)
    bare, dense = lt == 0.0, lt >= 2.0
    assert np.all(a["displa"][bare] == 0.0)
    np.testing.assert_allclose(a["z0mv"][bare], z0mg[bare], rtol=1e-14)
    np.testing.assert_array_equal(a["displa"][dense], d_in[dense])
    np.testing.assert_allclose(a["z0mv"][dense], z_in[dense], rtol=1e-14)
    np.testing.assert_array_equal(a["z0hv"], a["z0mv"])
    hu = st0.can["forc_hgt_u"][st0.can["gridcell"] - 1]
    np.testing.assert_allclose(a["forc_hgt_u_patch"], hu + a["z0mv"] + a["displa"], rtol=1e-15)


def test_meier_roughness_forms():
    from scipy.optimize import brentq
    before, st = _canopy_run(z0param_method=1)
    g, pft = st.can, S.pft_table()
    it = g["itype"]
    htop = g["htop"]
    lt = np.maximum(1e-5, g["elai"] + g["esai"])
    x = np.sqrt(7.5 * lt)
    d = htop * (1.0 - (1.0 - np.exp(-x)) / x)
    np.testing.assert_allclose(g["displa"], d, rtol=1e-13)
    ltm = np.minimum(lt, pft["z0v_LAImax"][it])
    z0 = np.empty_like(htop)
    for i in range(len(htop)):
        x0 = (pft["z0v_Cs"][it[i]] + pft["z0v_Cr"][it[i]] * ltm[i] * 0.5) ** -0.5 * pft["z0v_c"][it[i]] * \
            ltm[i] * 0.25
        u = brentq(lambda v: v - x0 * np.exp(v), 0.0, 1.0)       # smaller root of v = x0 e^v
        u = 4.0 * u / ltm[i] / pft["z0v_c"][it[i]]
        cw = pft["z0v_cw"][it[i]]
        z0[i] = htop[i] * (1.0 - d[i] / htop[i]) * np.exp(-0.4 * u + np.log(cw) - 1.0 + 1.0 / cw)
    # the reference stops its fixed-point iteration at |dx| < 1e-4
    np.testing.assert_allclose(g["z0mv"], z0, rtol=2e-3)


@pytest.mark.parametrize("biomass", [0, 1])
def test_resistances_closed_forms(biomass):
    before, st = _canopy_run(use_biomass_heat_storage=biomass, use_undercanopy_stability=0)
    g, prm, pft = st.can, S.canopy_params(), S.pft_table()
    ustar, uaf = g["ustar"], g["uaf"]
    np.testing.assert_allclose(uaf, ustar, rtol=1e-14)
    # the wind of the last iteration's resistances (um is updated after them
    # for the next iteration): r_am = u_m / u*^2
    um = g["ram1"] * ustar ** 2
    dleaf = pft["dleaf"][g["itype"]]
    np.testing.assert_allclose(g["rb1"], np.sqrt(dleaf / uaf) / prm["cv"], rtol=1e-13)
    z0mg = st.can["z0mg"][st.pch["column"] - 1]
    w = np.exp(-(g["elai"] + g["esai"]))
    cs_bare = 0.4 / (prm["a_coef"] * (z0mg * uaf / 1.5e-5) ** prm["a_exp"])
    cs = cs_bare * w + prm["csoilc"] * (1.0 - w)
    u_under = np.minimum(0.4, 0.03 * um / ustar) if biomass else uaf
    np.testing.assert_allclose(g["rah2"], 1.0 / (cs * u_under), rtol=1e-12)
    np.testing.assert_array_equal(g["raw2"], g["rah2"])


# ---------------------------------------------------------------------------
# QSat against the published Flatau et al. (1992) fits

FLATAU = {  # 8th-order relative-error fits, es in mb, t in C (Flatau et al. 1992, Table 4)
    "water": ["6.11213476", "0.444007856", "0.143064234e-01", "0.264461437e-03", "0.305903558e-05",
              "0.196237241e-07", "0.892344772e-10", "-0.373208410e-12", "0.209339997e-15"],
    "dwater": ["0.444017302", "0.286064092e-01", "0.794683137e-03", "0.121211669e-04", "0.103354611e-06",
               "0.404125005e-09", "-0.788037859e-12", "-0.114596802e-13", "0.381294516e-16"],
    "ice": ["6.11123516", "0.503109514", "0.188369801e-01", "0.420547422e-03", "0.614396778e-05",
            "0.602780717e-07", "0.387940929e-09", "0.149436277e-11", "0.262655803e-14"],
    "dice": ["0.503277922", "0.377289173e-01", "0.126801703e-02", "0.249468427e-04", "0.313703411e-06",
             "0.257180651e-08", "0.133268878e-10", "0.394116744e-13", "0.498070196e-16"],
}


def _poly(coef, td):
    """Exact value of the polynomial at td and the rounding bound of its
    Horner evaluation in binary64 (20 ulp of the sum of |terms|)."""
    t = Fraction(td)
    terms = [Fraction(a) * t ** k for k, a in enumerate(coef)]
    return float(sum(terms)), 20 * 2.0 ** -53 * float(sum(abs(x) for x in terms))


def test_qsat_flatau_table():
    T = np.concatenate([np.linspace(190.0, 380.0, 401), [TFRZ, TFRZ - 1e-9, 198.15 - 5, 373.15 + 5]])
    p = np.full_like(T, 85000.0)
    qs, es, dqs = O.qsat(T, p)
    for i, t in enumerate(T):
        td = min(100.0, max(-75.0, t - TFRZ))
        e, ebound = _poly(FLATAU["water" if td >= 0 else "ice"], td)
        de, debound = _poly(FLATAU["dwater" if td >= 0 else "dice"], td)
        e, ebound, de, debound = 100.0 * e, 100.0 * ebound, 100.0 * de, 100.0 * debound
        assert abs(es[i] - e) <= ebound, t
        vp = 1.0 / (p[i] - 0.378 * e)
        assert qs[i] == pytest.approx(0.622 * e * vp, rel=4e-15 + ebound / e), t
        assert dqs[i] == pytest.approx(de * 0.622 * vp * vp * p[i], rel=6e-15 + debound / de + 2 * ebound / e), t


# ---- EnergyBalanceCheck's patch errors (BalanceCheckMod.F90:966-1003) --------

@pytest.mark.parametrize("workload", ["canopy", "canopy_uneven"])
def test_energy_balance_errors_close_after_land_steps(workload):
    """errseb and errlon are energy residuals of the step: after full land
    steps (CanopyFluxes, SoilTemperature, SoilFluxes, hydrology) they close
    to round-off (the reference warns at 1e-7 W/m2), so a dropped term or a
    wrong sign in the check -- each term is tens to hundreds of W/m2 -- fails
    here; inactive patches report 0 and the max / maxloc are those of the
    arrays (lowest patch on ties)."""
    from ctsm_amd import synthetic as S
    st = S.canopy_workload(workload, ncol=192)
    st.can["active"][::7] = 0
    cfg, prm, pft = O.canopy_config(), S.canopy_params(), O.PftStruct(S.pft_table())
    for _ in range(2):
        rc, bad = O.land_step_full(st, cfg, prm, pft, hydro=S.hydro_params(), water_table_after=True,
                                   drainage_after=True)
        assert rc == 0, bad
    seb, lon, (ms, ps), (ml, pl) = O.energy_balance(st)
    act = st.can["active"] != 0
    assert np.all(seb[~act] == 0.0) and np.all(lon[~act] == 0.0)
    assert np.max(np.abs(seb)) < 1e-9 and np.max(np.abs(lon)) < 1e-11
    assert ms == np.max(np.abs(seb)) and ps == int(np.argmax(np.abs(seb))) + 1
    assert ml == np.max(np.abs(lon)) and pl == int(np.argmax(np.abs(lon))) + 1
    # the terms are not small: the closure is a real check
    assert np.max(np.abs(st.sf["eflx_soil_grnd"][act])) > 10.0 and np.max(np.abs(st.can["dhsdt_canopy"][act])) > 1e-3
