"""Pins of the oracle's physics restatement (CPU only).

Nothing in the reference's own test suite pins SoilWater / SoilTemperature
(no pFUnit test exists for them), so the oracle is pinned by what the
mathematics and the reference's runtime checks fix:

  Richards (SoilWaterMovementMod.F90:976-1440)
    - Clapp-Hornberger closed forms at saturation and half saturation, ice
      impedance 10^(-e_ice*icefrac);
    - the flux derivatives that form the implicit matrix equal finite
      differences of the fluxes (a wrong sign or index fails this);
    - a hydrostatic column (total head constant) with no sources is stationary
      (pins the sign of gravity and of the suction term);
    - column water conservation: sum of the change = (infiltration - root
      sink) * dtime with a zero-flux bottom (pins the flux bookkeeping);
    - the sub-step schedule: tiny fluxes give one sub-step.
  Heat solve + phase change (SoilTemperatureMod.F90:92-1540)
    - the reference's own soil energy balance check errsoi
      (SoilFluxesMod.F90:399-425, abort threshold 1e-4 W/m2) closes to
      round-off with and without snow and through heavy phase change;
    - an isothermal column with zero net surface flux stays isothermal;
    - liquid + ice mass of every layer is conserved by the phase change;
    - the Sturm (1997) and Jordan (1991) snow conductivities equal the
      published formulas.
"""
import numpy as np
import pytest

from ctsm_amd import synthetic
from oracle import oracle as O

NS = 12


def _uniform_column(ncol=1, s=0.6, sand=40.0, clay=30.0):
    st = synthetic.workload("cpu1024", ncol=ncol)
    prm = synthetic.soil_params(np.full((25, ncol), sand), np.full((25, ncol), clay), np.zeros((25, ncol)),
                                synthetic.vertical_grid()[2][:, None])
    for k, v in prm.items():
        st.col[k][...] = v
    dz = st.col["dz"][NS:NS + 25]
    st.col["h2osoi_liq"][NS:NS + 20] = (s * st.col["watsat"][:20] * dz[:20] * 1000.0)
    st.col["icefrac"][...] = 0.0
    st.col["eff_porosity"][...] = st.col["watsat"]
    return st


def test_ch78_closed_forms():
    st = _uniform_column(s=1.0)
    sysr = O.richards_system(st, 1, 1800.0)
    ws, b = st.col["watsat"][0, 0], st.col["bsw"][0, 0]
    assert np.allclose(sysr["smp"], -st.col["sucsat"][:20, 0], rtol=1e-12)    # smp(s=1) = -sucsat
    assert np.allclose(sysr["hk"], st.col["hksat"][:20, 0], rtol=1e-12)       # hk(s=1) = imped*hksat, imped=1
    st = _uniform_column(s=0.5)
    sysr = O.richards_system(st, 1, 1800.0)
    assert np.allclose(sysr["smp"], -st.col["sucsat"][:20, 0] * 2.0 ** b, rtol=1e-12)
    assert np.allclose(sysr["hk"], st.col["hksat"][:20, 0] * 0.5 ** (2 * b + 3), rtol=1e-12)
    assert np.allclose(sysr["dhkdw"], (2 * b + 3) * sysr["hk"] / 0.5, rtol=1e-12)
    st.col["icefrac"][...] = 0.5
    sysr = O.richards_system(st, 1, 1800.0)
    assert np.allclose(sysr["imped"], 10.0 ** (-6.0 * 0.5), rtol=1e-13)


def test_flux_derivatives_match_finite_differences():
    rng = np.random.default_rng(3)
    st = _uniform_column(s=0.5)
    dz = st.col["dz"][NS:NS + 20, 0]
    ws = st.col["watsat"][:20, 0]
    # a non-uniform, unclamped profile (0.3 < s < 0.8 everywhere)
    s = rng.uniform(0.3, 0.8, 20)
    st.col["h2osoi_liq"][NS:NS + 20, 0] = s * ws * dz * 1000.0
    base = O.richards_system(st, 1, 1800.0)
    n = 20
    for k in range(n - 1):           # perturb layer k, watch qout(k) and qout(k-1)
        h = 1e-6 * s[k] * ws[k]
        vals = []
        for sgn in (+1, -1):
            st2 = st.copy()
            st2.col["h2osoi_liq"][NS + k, 0] += sgn * h * dz[k] * 1000.0
            vals.append(O.richards_system(st2, 1, 1800.0)["qout"])
        fd = (vals[0] - vals[1]) / (2 * h)
        # d qout(k) / d vwc(k) = dqodw1(k);  d qout(k-1) / d vwc(k) = dqodw2(k-1)
        assert np.isclose(fd[k], base["dqodw1"][k], rtol=1e-5, atol=1e-12 * abs(base["qout"]).max() / h), k
        if k >= 1:
            assert np.isclose(fd[k - 1], base["dqodw2"][k - 1], rtol=1e-5,
                              atol=1e-12 * abs(base["qout"]).max() / h), k
        # qin(k+1) is qout(k) with the same derivatives, and the LHS uses them
        assert base["qin"][k + 1] == base["qout"][k]
        assert base["dqidw0"][k + 1] == base["dqodw1"][k]


def test_hydrostatic_column_is_stationary():
    st = _uniform_column(s=0.5)
    suc, b = st.col["sucsat"][0, 0], st.col["bsw"][0, 0]
    ws = st.col["watsat"][0, 0]
    z = st.col["z"][NS:NS + 20, 0]
    dz = st.col["dz"][NS:NS + 20, 0]
    smp1 = -suc * 0.3 ** (-b)
    smp = smp1 + 1000.0 * (z - z[0])           # total head smp - 1000 z constant
    s = (-smp / suc) ** (-1.0 / b)
    assert s.max() < 1.0
    st.col["h2osoi_liq"][NS:NS + 20, 0] = s * ws * dz * 1000.0
    st.col["qflx_infl"][:] = 0.0
    st.col["qflx_rootsoi"][...] = 0.0
    before = st.col["h2osoi_liq"][NS:NS + 20, 0].copy()
    # the interface conductivity uses the mean saturation, so the equilibrium
    # flux is zero only to second order in the layer spacing
    sysr = O.richards_system(st, 1, 1800.0)
    assert np.max(np.abs(sysr["qout"])) < 1e-3 * st.col["hksat"][0, 0]
    O.soilwater(st, filter_=[1])
    after = st.col["h2osoi_liq"][NS:NS + 20, 0]
    assert np.max(np.abs(after - before) / before) < 1e-3
    assert st.col["num_substeps"][0] == 1


def test_richards_mass_conservation():
    st = synthetic.workload("uneven", ncol=2048)
    s0 = st.copy()
    O.soilwater(st)
    nb = st.col["nbedrock"]
    j = np.arange(25)[:, None] + 1
    m = j <= nb[None, :]
    w0 = np.where(m, s0.col["h2osoi_liq"][NS:NS + 25], 0).sum(0)
    w1 = np.where(m, st.col["h2osoi_liq"][NS:NS + 25], 0).sum(0)
    sink = np.where(m, st.col["qflx_rootsoi"], 0).sum(0)
    err = (w1 - w0) - 1800.0 * (st.col["qflx_infl"] - sink)
    # relative to the column's water: the implicit solve conserves to round-off
    assert np.max(np.abs(err) / w0) < 1e-7
    assert np.median(np.abs(err)) < 1e-9
    assert np.all(st.col["num_substeps"] >= 1)


def test_tiny_fluxes_take_one_substep():
    st = _uniform_column(ncol=4, s=0.5)
    st.col["qflx_infl"][:] = 1e-12
    st.col["qflx_rootsoi"][...] = 0.0
    O.soilwater(st)
    assert np.all(st.col["num_substeps"] == 1)


def _errsoi_run(st, nsteps):
    worst = 0.0
    for _ in range(nsteps):
        tss, thb, tg0 = st.col["t_soisno"].copy(), st.col["t_h2osfc"].copy(), st.col["t_grnd"].copy()
        rc, _ = O.soiltemperature(st)
        assert rc == 0
        e = O.errsoi(st, tss, thb, tg0)
        worst = max(worst, float(np.max(np.abs(e))))
        O.soilwater_fractions(st)
        O.soilwater(st)
    return worst


@pytest.mark.parametrize("name", ["cpu1024", "uneven"])
def test_errsoi_closes(name):
    st = synthetic.workload(name, ncol=1024)
    worst = _errsoi_run(st, 4)
    assert worst < 1e-7, worst            # reference warns at 1e-5, aborts at 1e-4 W/m2


def test_errsoi_closes_with_standing_water():
    st = synthetic.soil_only(512, 77, snow=True, bedrock=True, frac_h2osfc=np.full(512, 0.3))
    st.col["t_h2osfc"][:] = 265.0
    synthetic.set_t_grnd_consistent(st)
    worst = _errsoi_run(st, 3)
    assert worst < 1e-6, worst


def test_isothermal_zero_flux_column_stays_isothermal():
    st = synthetic.workload("uneven", ncol=64)
    T = 280.0
    st.col["t_soisno"][...] = T
    st.col["h2osoi_ice"][NS:] = 0.0                          # no soil ice to melt
    st.col["snl"][:] = 0
    st.col["h2osoi_ice"][:NS] = 0.0
    st.col["frac_sno_eff"][:] = 0.0
    st.col["t_grnd"][:] = T
    st.col["t_h2osfc"][:] = T
    st.col["forc_lwrad"][:] = 5.67e-8 * T ** 4
    st.pch["sabg"][:] = st.pch["sabg_soil"][:] = st.pch["sabg_snow"][:] = 0.0
    st.pch["sabg_lyr"][...] = 0.0
    O.soiltemperature(st)
    assert np.max(np.abs(st.col["t_soisno"][NS:] - T)) < 1e-9
    assert np.all(st.col["imelt"][NS:] == 0)


def test_phase_change_conserves_layer_mass():
    st = synthetic.workload("uneven", ncol=2048)
    w0 = st.col["h2osoi_liq"] + st.col["h2osoi_ice"]
    O.soiltemperature(st)
    w1 = st.col["h2osoi_liq"] + st.col["h2osoi_ice"]
    assert np.allclose(w0, w1, rtol=1e-13, atol=1e-12)
    assert np.any(st.col["imelt"] == 2) and np.any(st.col["xmf"] != 0)


@pytest.mark.parametrize("method", [0, 1])
def test_snow_conductivity_published_forms(method):
    st = synthetic.workload("uneven", ncol=256)
    O.soiltemperature(st, cfg=O.config(snow_thermal_cond_method=method))
    snl = st.col["snl"]
    rows = np.arange(NS)[:, None] - NS + 1     # level of each snow row
    act = rows >= snl[None, :] + 1
    bw = st.col["bw"][act]
    thk = st.col["thk"][:NS][act]
    assert bw.size > 100
    if method == 0:       # Jordan (1991): tkair + (7.75e-5 rho + 1.105e-6 rho^2)(tkice - tkair)
        ref = 0.023 + (7.75e-5 * bw + 1.105e-6 * bw ** 2) * (2.29 - 0.023)
        assert np.allclose(thk, ref, rtol=1e-13)
    else:                 # Sturm (1997), rho in g/cm3: 0.023 + 0.234 rho | 0.138 - 1.01 rho + 3.233 rho^2
        rho = bw / 1000.0
        ref = np.where(bw <= 156, 0.023 + 0.234 * rho, 0.138 - 1.01 * rho + 3.233 * rho ** 2)
        assert np.allclose(thk, ref, rtol=2e-7)          # single-precision literals in the reference
