"""The reference's own acceptance checks applied to the GPU results.

CTSM validates every step on the host (BalanceCheckMod.F90): the soil energy
residual errsoi (SoilFluxesMod.F90:399-445; warning above 1e-5 W/m2, abort
above 1e-4 W/m2, BalanceCheckMod.F90:1103-1121), the column water balance
(errh2o, BalanceCheckMod.F90:617-630: warning above 1e-9 mm, abort above
1e-5 mm; here the Richards solve's own part, the column's liquid water against
infiltration and the root sink with a zero-flux bottom), and CanopyFluxes'
leaf energy residual err <= 0.1 W/m2 (CanopyFluxesMod.F90:1742-1756).  These
run the HIP step and apply the checks to what it returns, step by step, over
a 48-step day.  The bar for errh2o is "no larger than the reference's": the
implicit solve conserves water only to the round-off of its linear systems,
and a column of ~3,000 mm taking 35 sub-steps leaves ~2e-5 mm in the oracle
too, so the GPU's worst error is held to the oracle's worst on the same
steps (or 1e-5 mm, whichever is larger)."""
import numpy as np
import pytest

from ctsm_amd import synthetic as S
from ctsm_amd.state import DeviceState
from oracle import oracle as O
from tests.helpers import NS

pytestmark = pytest.mark.gpu


def _water(st):
    nb = st.col["nbedrock"]
    m = (np.arange(st.col["watsat"].shape[0])[:, None] + 1) <= nb[None, :]
    liq = np.where(m, st.col["h2osoi_liq"][NS:NS + m.shape[0]], 0.0).sum(0)
    sink = np.where(m, st.col["qflx_rootsoi"], 0.0).sum(0)
    return liq, sink


def _oracle_worst_errh2o(st, nsteps):
    st = st.copy()
    worst = 0.0
    for _ in range(nsteps):
        rc, _ = O.soiltemperature(st)
        assert rc == 0
        w0, _ = _water(st)
        O.soilwater_fractions(st)
        rc, _ = O.soilwater(st)
        assert rc == 0
        w1, sink = _water(st)
        worst = max(worst, float(np.max(np.abs((w1 - w0) - 1800.0 * (st.col["qflx_infl"] - sink)))))
    return worst


@pytest.mark.parametrize("name,ncol", [("uneven", 4096), ("cpu1024", 1024)])
def test_soil_step_balance_checks_over_a_day(ctx, name, ncol):
    st = S.workload(name, ncol=ncol)
    ref_h2o = _oracle_worst_errh2o(st, 48)
    print(f"case: balance checks of the fused soil step, {name}, {ncol} columns, 48 steps", flush=True)
    ds = DeviceState(st, ctx)
    worst_soi = worst_h2o = 0.0
    try:
        cur = ds.to_host()
        for step in range(48):
            tss, thb, tg0 = cur.col["t_soisno"].copy(), cur.col["t_h2osfc"].copy(), cur.col["t_grnd"].copy()
            # SoilTemperature alone first, so errsoi sees the heat solve's own state
            ctx.soiltemperature_dev(ds)
            ctx.sync()
            mid = ds.to_host()
            e = O.errsoi(mid, tss, thb, tg0)
            worst_soi = max(worst_soi, float(np.max(np.abs(e))))
            w0, _ = _water(mid)
            ctx.soilwater_fractions_dev(ds)
            ctx.soilwater_dev(ds)
            ctx.sync()
            cur = ds.to_host()
            w1, sink = _water(cur)
            err = (w1 - w0) - 1800.0 * (cur.col["qflx_infl"] - sink)
            worst_h2o = max(worst_h2o, float(np.max(np.abs(err))))
    finally:
        ds.free()
    print(f"max |errsoi| {worst_soi:.3e} W/m2, max |errh2o| {worst_h2o:.3e} mm (oracle {ref_h2o:.3e} mm)",
          flush=True)
    assert worst_soi < 1e-5, worst_soi       # below even the reference's warning level
    assert worst_h2o <= max(1e-5, ref_h2o * (1.0 + 1e-6)), (worst_h2o, ref_h2o)


def test_land_step_canopy_energy_residual(ctx):
    ctx.set_canopy_params(S.canopy_params(), S.pft_table())
    st = S.canopy_workload("canopy_uneven", ncol=2048)
    print(f"case: canopy energy residual over 6 land steps, 2,048 columns / {st.dims.npatch} patches", flush=True)
    ds = DeviceState(st, ctx)
    worst = 0.0
    try:
        for _ in range(6):
            ctx.land_step_dev(ds)
            ctx.sync()
            g = ds.to_host()
            for k, v in g.can.items():
                assert np.all(np.isfinite(v)), k
            worst = max(worst, float(np.max(np.abs(g.can["canopy_err"]))))
            assert np.all(g.can["num_iter"] >= 3) and np.all(g.can["num_iter"] <= 41)
    finally:
        ds.free()
    print(f"max |canopy err| {worst:.3e} W/m2", flush=True)
    assert worst <= 0.1
