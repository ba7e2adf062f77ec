"""Seeded synthetic soil columns for the benchmark configurations.

Every random number is a pure function of (seed, stream, global column id,
level), computed by a splitmix64-style integer hash, so any column can be
regenerated alone (the served-ensemble determinism test relies on this) and
the generator vectorises over a million columns.

Soil parameters follow the reference's initialisation path:
  - 20SL_8.5m layer thicknesses, initVerticalMod.F90:291-303 and :314-319
  - Cosby et al. (1984) table 5 pedotransfer, FuncPedotransferMod.F90:97-101
  - organic-matter mixing and thermal properties,
    SoilStateInitTimeConstMod.F90:540-589 (param-file scalars from the CLM5
    technical note: tkd_sand 8.80, tkd_clay 2.92, tkm_om 0.25, tkd_om 0.05,
    csol_sand 2.128e6, csol_clay 2.385e6, csol_om 2.5e6, pd 2700, organic_max
    130 kg/m3, zsapric 0.5 m, pcalpha 0.5, pcbeta 0.139; scale factors 1).

Workloads (sizes are the caller's):
  soil_only(..., snow=False, bedrock=False): 1,024-column CPU case -- no
      snow, 20 Richards layers.
  soil_only(..., snow=True, bedrock=True): uneven columns -- 0..12 snow
      layers (thickness U(0.01,0.5) m, density U(50,500) kg/m3, T U(240,273) K)
      and bedrock index U{5..20}.

The ground heat flux G ~ U(-100,200) W/m2 enters as one patch per column
(wtcol 1): sabg = sabg_soil = sabg_snow = G absorbed at the top layer,
downward longwave equal to the black-body flux of the initial ground
temperature (so the net longwave term emg*(forc_lwrad - sb*T^4) starts at
zero and damps drift), emg 0.96, no sensible or latent flux, cgrnd 0.  The
soil therefore sees G plus a longwave restoring term with
dhsdT = -4 emg sb T_grnd^3.
"""
import numpy as np

from .state import Dims, HostState

SB = 5.67e-8
HVAP = 2.501e6
TFRZ = 273.15

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def _mix(x):
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * _M1
        x = (x ^ (x >> np.uint64(27))) * _M2
        return x ^ (x >> np.uint64(31))


def uniform(seed, stream, cols, k=0):
    """U[0,1) for global column ids `cols` (int array), level/draw index k."""
    with np.errstate(over="ignore"):
        key = _mix(np.uint64(seed) * _G + np.uint64(stream) * np.uint64(0xD1B54A32D192ED03)
                   + np.uint64(k) * np.uint64(0x8CB92BA72F3D8DD7) + np.uint64(1))
        x = _mix(np.asarray(cols, dtype=np.uint64) * _G + key)
        x = _mix(x ^ key)
    return (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def vertical_grid(nlevsoi=20, nlevgrnd=25):
    """20SL_8.5m dzsoi, zisoi(0:nlevgrnd), zsoi (initVerticalMod.F90:291-319)."""
    dz = np.zeros(nlevgrnd + 1)
    for j in range(1, 5):
        dz[j] = j * 0.02
    for j in range(5, 14):
        dz[j] = dz[4] + (j - 4) * 0.04
    for j in range(14, nlevsoi + 1):
        dz[j] = dz[13] + (j - 13) * 0.10
    for j in range(nlevsoi + 1, nlevgrnd + 1):
        dz[j] = dz[nlevsoi] + (((j - nlevsoi) * 25.0) ** 1.5) / 100.0
    zi = np.zeros(nlevgrnd + 1)
    for j in range(1, nlevgrnd + 1):
        zi[j] = sum(dz[1:j + 1])
    z = np.zeros(nlevgrnd + 1)
    for j in range(1, nlevgrnd + 1):
        z[j] = 0.5 * (zi[j - 1] + zi[j])
    return dz[1:], zi, z[1:]


# CLM5 technical-note values of the param-file scalars used by the mixing.
TKD_SAND, TKD_CLAY, TKM_OM, TKD_OM = 8.80, 2.92, 0.25, 0.05
CSOL_SAND, CSOL_CLAY, CSOL_OM = 2.128, 2.385, 2.5
PD, ORGANIC_MAX, ZSAPRIC, PCALPHA, PCBETA = 2700.0, 130.0, 0.5, 0.5, 0.139


def soil_params(sand, clay, om, zsoi, sf=None):
    """Hydraulic and thermal parameters for arrays of sand %, clay %, organic
    matter (kg/m3) at node depths zsoi (SoilStateInitTimeConstMod.F90:540-589).
    `sf` optional dict of scale factors (bsw_sf, hksat_sf, sucsat_sf, watsat_sf)."""
    sf = sf or {}
    om_frac = np.minimum(om / ORGANIC_MAX, 1.0)
    watsat = 0.489 - 0.00126 * sand
    sucsat = 10.0 * (10.0 ** (1.88 - 0.0131 * sand))
    xksat = 0.0070556 * (10.0 ** (-0.884 + 0.0153 * sand))
    om_watsat = np.maximum(0.93 - 0.1 * (zsoi / ZSAPRIC), 0.83)
    om_b = np.minimum(2.7 + 9.3 * (zsoi / ZSAPRIC), 12.0)
    om_sucsat = np.minimum(10.3 - 0.2 * (zsoi / ZSAPRIC), 10.1)
    om_hksat = np.maximum(0.28 - 0.2799 * (zsoi / ZSAPRIC), xksat)
    bd = (1.0 - watsat) * PD
    watsat = np.minimum(sf.get("watsat_sf", 1.0) * ((1.0 - om_frac) * watsat + om_watsat * om_frac), 0.93)
    tkm = (1.0 - om_frac) * (TKD_SAND * sand + TKD_CLAY * clay) / (sand + clay) + TKM_OM * om_frac
    bsw = sf.get("bsw_sf", 1.0) * ((1.0 - om_frac) * (2.91 + 0.159 * clay) + om_frac * om_b)
    sucsat = sf.get("sucsat_sf", 1.0) * ((1.0 - om_frac) * sucsat + om_sucsat * om_frac)
    perc_norm = (1.0 - PCALPHA) ** (-PCBETA)
    perc_frac = np.where(om_frac > PCALPHA,
                         perc_norm * np.maximum(om_frac - PCALPHA, 0.0) ** PCBETA, 0.0)
    uncon_frac = (1.0 - om_frac) + (1.0 - perc_frac) * om_frac
    with np.errstate(divide="ignore", invalid="ignore"):
        uncon_hksat = np.where(om_frac < 1.0,
                               uncon_frac / ((1.0 - om_frac) / xksat + ((1.0 - perc_frac) * om_frac) / om_hksat),
                               0.0)
    hksat = sf.get("hksat_sf", 1.0) * (uncon_frac * uncon_hksat + (perc_frac * om_frac) * om_hksat)
    tkmg = tkm ** (1.0 - watsat)
    tkdry = ((0.135 * bd + 64.7) / (PD - 0.947 * bd)) * (1.0 - om_frac) + TKD_OM * om_frac
    csol = ((1.0 - om_frac) * (CSOL_SAND * sand + CSOL_CLAY * clay) / (sand + clay) + CSOL_OM * om_frac) * 1.0e6
    return dict(watsat=watsat, bsw=bsw, sucsat=sucsat, hksat=hksat, tkmg=tkmg, tkdry=tkdry, csol=csol)


# RNG stream ids (one per drawn quantity)
S_SAND, S_CLAY, S_OM, S_LIQ, S_T, S_INFL, S_SINKTOT, S_SINKW, S_G, S_NSNO, S_SDZ, S_SRHO, S_ST, S_BED, S_ZWTC = range(15)


def _gids(ncol, col0, gids):
    """Global ids of the generated columns: col0 .. col0 + ncol - 1, or the
    given ids (any order: a binned order, a sample) when gids is not None."""
    if gids is None:
        return np.arange(col0, col0 + ncol, dtype=np.int64)
    gids = np.asarray(gids, dtype=np.int64)
    assert gids.shape == (ncol,)
    return gids


def structure(n, seed, snow=True, bedrock=True, nlevsno=12, nlevsoi=20):
    """(snow layers, bedrock index) of global columns 0 .. n-1, as soil_only
    draws them -- enough to bin a batch before generating it."""
    gid = np.arange(n, dtype=np.int64)
    nbed = (5 + np.floor(uniform(seed, S_BED, gid) * 16).astype(np.int32)) if bedrock \
        else np.full(n, nlevsoi, np.int32)
    nsno = np.floor(uniform(seed, S_NSNO, gid) * (nlevsno + 1)).astype(np.int32) if snow \
        else np.zeros(n, np.int32)
    return np.minimum(nsno, nlevsno), np.minimum(nbed, nlevsoi)


def soil_only(ncol, seed, snow=False, bedrock=False, col0=0, nlevsno=12, nlevsoi=20, nlevgrnd=25,
              frac_h2osfc=None, gids=None):
    """Columns of the soil-only workloads (one patch per column).

    col0 is the global id of the first column (for regenerating a slice);
    gids, when given, lists the global ids of the columns to generate."""
    d = Dims(ncol, ncol, nlevsno, nlevsoi, nlevgrnd)
    st = HostState(d)
    c, p = st.col, st.pch
    gid = _gids(ncol, col0, gids)
    dzs, zis, zs = vertical_grid(nlevsoi, nlevgrnd)
    NS = nlevsno

    # ---- column structure
    nbed = (5 + np.floor(uniform(seed, S_BED, gid) * 16).astype(np.int32)) if bedrock \
        else np.full(ncol, nlevsoi, np.int32)
    nbed = np.minimum(nbed, nlevsoi)
    nsno = np.floor(uniform(seed, S_NSNO, gid) * (nlevsno + 1)).astype(np.int32) if snow \
        else np.zeros(ncol, np.int32)
    nsno = np.minimum(nsno, nlevsno)
    c["snl"][:] = -nsno
    c["nbedrock"][:] = nbed
    c["patchi"][:] = np.arange(1, ncol + 1, dtype=np.int32)
    c["patchf"][:] = c["patchi"]

    # geometry: soil levels are rows NS .. NS+nlevgrnd-1 of the sno arrays
    for k in (c["dz"], c["z"], c["zi"]):
        k[...] = 0.0
    c["dz"][NS:, :] = dzs[:, None]
    c["z"][NS:, :] = zs[:, None]
    c["zi"][NS:, :] = zis[:, None]            # zi rows: j=-nlevsno at row 0, j=0 at row NS
    # snow layers j = 0, -1, ..., snl+1 (row NS-1+j in sno arrays; zi row NS+j)
    liq, ice, t = c["h2osoi_liq"], c["h2osoi_ice"], c["t_soisno"]
    liq[...] = 0.0
    ice[...] = 0.0
    t[...] = 0.0
    ztop = np.zeros(ncol)
    snow_w = np.zeros(ncol)
    for jj in range(0, nlevsno):           # jj = -j: 0 is the bottom snow layer
        j = -jj
        have = nsno > jj
        sdz = 0.01 + 0.49 * uniform(seed, S_SDZ, gid, jj)
        rho = 50.0 + 450.0 * uniform(seed, S_SRHO, gid, jj)
        ts = 240.0 + 33.0 * uniform(seed, S_ST, gid, jj)
        r = NS - 1 + j
        c["dz"][r] = np.where(have, sdz, 0.0)
        c["z"][r] = np.where(have, ztop - 0.5 * sdz, 0.0)
        c["zi"][NS + j - 1] = np.where(have, ztop - sdz, 0.0)
        ice[r] = np.where(have, rho * sdz, 0.0)
        t[r] = np.where(have, ts, 0.0)
        snow_w += np.where(have, rho * sdz, 0.0)
        ztop = np.where(have, ztop - sdz, ztop)

    # ---- soil texture and parameters (per column and layer)
    for name in ("watsat", "bsw", "sucsat", "hksat", "tkmg", "tkdry", "csol"):
        c[name][...] = 0.0
    sand_d = np.zeros((nlevgrnd, ncol))
    clay_d = np.zeros((nlevgrnd, ncol))
    om_d = np.zeros((nlevgrnd, ncol))
    for j in range(1, nlevsoi + 1):
        sa = 5.0 + 90.0 * uniform(seed, S_SAND, gid, j)
        cl = 5.0 + 90.0 * uniform(seed, S_CLAY, gid, j)
        tot = sa + cl
        scale = np.where(tot > 100.0, 100.0 / tot, 1.0)
        sand_d[j - 1], clay_d[j - 1] = sa * scale, cl * scale
        om_d[j - 1] = 130.0 * uniform(seed, S_OM, gid, j)
    for j in range(nlevsoi + 1, nlevgrnd + 1):    # below nlevsoi: deepest texture, no organic
        sand_d[j - 1], clay_d[j - 1] = sand_d[nlevsoi - 1], clay_d[nlevsoi - 1]
    prm = soil_params(sand_d, clay_d, om_d, zs[:, None])
    for k, v in prm.items():
        c[k][...] = v

    # ---- soil water and temperature
    for j in range(1, nlevgrnd + 1):
        r = NS - 1 + j
        u = uniform(seed, S_LIQ, gid, j)
        w = (0.1 + 0.85 * u) * c["watsat"][j - 1] * dzs[j - 1] * 1000.0
        liq[r] = np.where(j <= nbed, w, 0.0)
        ice[r] = 0.0
        t[r] = 260.0 + 40.0 * uniform(seed, S_T, gid, j)

    # ---- fluxes into the soil-water solve
    c["qflx_infl"][:] = 1.0e-3 * uniform(seed, S_INFL, gid)
    tot = 1.0e-4 * uniform(seed, S_SINKTOT, gid)
    wts = np.zeros((nlevgrnd, ncol))
    for j in range(1, nlevsoi + 1):
        e = -np.log(1.0 - uniform(seed, S_SINKW, gid, j))
        wts[j - 1] = np.where(j <= nbed, e, 0.0)
    wts /= wts.sum(axis=0, keepdims=True)
    c["qflx_rootsoi"][...] = wts * tot[None, :]

    # ---- snow / surface-water scalars
    fh = np.zeros(ncol) if frac_h2osfc is None else np.asarray(frac_h2osfc, dtype=np.float64)
    # snow and standing water share the ground: frac_sno_eff + frac_h2osfc <= 1
    c["frac_sno_eff"][:] = np.where(nsno > 0, 1.0 - fh, 0.0)
    c["frac_h2osfc"][:] = fh
    c["h2osfc"][:] = np.where(fh > 0, 2.0, 0.0)
    c["h2osno_no_layers"][:] = 0.0
    c["int_snow"][:] = snow_w
    c["snow_depth"][:] = -ztop
    c["snomelt_accum"][:] = 0.0
    top_t = np.where(nsno > 0, t[NS - nsno, np.arange(ncol)], t[NS])
    c["t_h2osfc"][:] = t[NS]
    set_t_grnd_consistent(st)
    c["emg"][:] = 0.96
    c["htvp"][:] = HVAP
    c["eflx_bot"][:] = 0.0
    c["forc_lwrad"][:] = SB * top_t ** 4
    # read by SoilWater only with the water-table bottom boundary
    c["zwt"][:] = 0.2 + 7.8 * uniform(seed, S_ZWTC, gid)
    c["smpmin"][:] = -1.e8                          # SoilStateType's smpmin (mm)

    # soil-water fractions consistent with the initial ice (SoilHydrologyMod.F90:240-252)
    vol_ice = np.minimum(c["watsat"][:nlevgrnd], ice[NS:] / (c["dz"][NS:] * 917.0))
    c["eff_porosity"][...] = np.maximum(0.01, c["watsat"] - vol_ice)
    with np.errstate(invalid="ignore", divide="ignore"):
        c["icefrac"][...] = np.where(c["watsat"] > 0, np.minimum(1.0, vol_ice / c["watsat"]), 0.0)

    # ---- one patch per column carrying G
    G = -100.0 + 300.0 * uniform(seed, S_G, gid)
    p["column"][:] = np.arange(1, ncol + 1, dtype=np.int32)
    p["wtcol"][:] = 1.0
    p["frac_veg_nosno"][:] = 0
    for k in ("sabg", "sabg_soil", "sabg_snow"):
        p[k][:] = G
    p["sabg_lyr"][...] = 0.0
    p["sabg_lyr"][NS - nsno, np.arange(ncol)] = G       # row of level snl+1
    for k in ("dlrad", "cgrnd", "eflx_sh_grnd", "eflx_sh_snow", "eflx_sh_soil", "eflx_sh_h2osfc",
              "qflx_evap_soi", "qflx_ev_snow", "qflx_ev_soil", "qflx_ev_h2osfc"):
        p[k][:] = 0.0
    return st


def set_t_grnd_consistent(st):
    """t_grnd as SoilTemperature leaves it (SoilTemperatureMod.F90:457-477):
    the snow / soil / standing-water blend of the current temperatures."""
    c = st.col
    NS = st.dims.nlevsno
    snl = c["snl"]
    fs, fh = c["frac_sno_eff"], c["frac_h2osfc"]
    ncol = st.dims.ncol
    t1 = c["t_soisno"][NS]
    ttop = c["t_soisno"][NS + snl, np.arange(ncol)]
    th = c["t_h2osfc"]
    c["t_grnd"][:] = np.where(snl < 0, fs * ttop + (1.0 - fs - fh) * t1 + fh * th, (1.0 - fh) * t1 + fh * th)


def workload(name, ncol=None, seed=None):
    """Named workloads: 'cpu1024' (no snow, 20 layers, 1 step) and
    'uneven' (snow 0..12, bedrock 5..20, 48 steps)."""
    if name == "cpu1024":
        return soil_only(ncol or 1024, 0xC75D0000 if seed is None else seed)
    if name == "uneven":
        return soil_only(ncol or 65536, 0xC75D0001 if seed is None else seed, snow=True, bedrock=True)
    raise KeyError(name)


# ---------------------------------------------------------------------------
# Canopy workloads (1-16 vegetated patches per column)

S_NPATCH, S_PFT, S_LAI, S_SAI, S_HTOP, S_TV = 40, 41, 42, 43, 44, 45
S_FT, S_RH, S_WIND, S_WDIR, S_SW, S_LW, S_PBOT = 50, 51, 52, 53, 54, 55, 56
S_VCS, S_VCH, S_FWET, S_LIQCAN, S_T10, S_DAYL, S_NOON, S_BETA, S_RESIS = 60, 61, 62, 63, 64, 65, 66, 67, 68

CPAIR, RAIR = 1.00464e3, 6.02214e26 * 1.38065e-23 / 28.966


def canopy_params():
    """Representative CLM5/6 param-file scalars (the param file itself is not
    available offline; values follow the CLM5 technical note)."""
    return dict(cv=0.01, a_coef=0.13, a_exp=0.45, csoilc=0.004, z_dl=0.05, lai_dl=0.5, wind_min=1.0,
                zetamaxstable=2.0, kc25_coef=404.9e-6, ko25_coef=278.4e-3, cp25_yr2000=42.75e-6,
                kcha=79430.0, koha=36380.0, cpha=37830.0, vcmaxha=72000.0, jmaxha=50000.0, tpuha=72000.0,
                lmrha=46390.0, vcmaxhd=200000.0, jmaxhd=200000.0, tpuhd=200000.0, lmrhd=150650.0,
                lmrse=490.0, fnr=7.16, act25=60.0, theta_psii=0.7, theta_ip=0.95, fnps=0.15,
                tpu25ratio=0.167, kp25ratio=20000.0, vcmaxse_sf=1.0, jmaxse_sf=1.0, tpuse_sf=1.0,
                jmax25top_sf=1.0, leaf_mr_vcm=0.015)


def hydro_params(fff=0.5, pc=0.4, mu=0.14, h2osfcflag=1, aq_sp_yield_min=0.01, n_baseflow=1.0,
                 baseflow_scalar=5.5e-3, perched_baseflow_scalar=1.0e-5):
    """Representative param-file scalars of the hydrology (fff of
    SaturatedExcessRunoffMod, pc and mu of SurfaceWaterMod, the drainage
    parameters of SoilHydrologyMod; the param file itself is not available
    offline) and h2osfcflag."""
    from .abi import HydroParams
    return HydroParams(fff=fff, pc=pc, mu=mu, aq_sp_yield_min=aq_sp_yield_min, n_baseflow=n_baseflow,
                       baseflow_scalar=baseflow_scalar, perched_baseflow_scalar=perched_baseflow_scalar,
                       h2osfcflag=h2osfcflag, pad=0)


def pft_table():
    """Rows 0..16 (bare soil, 14 natural PFTs, two crops) with representative
    pftcon values; row 14 is the C4 grass."""
    n = 17
    trees = np.zeros(n, np.int32)
    trees[1:9] = 1
    shrubs = np.zeros(n, np.int32)
    shrubs[9:12] = 1
    t = dict(
        dleaf=np.full(n, 0.04),
        c3psn=np.where(np.arange(n) == 14, 0.0, 1.0),
        slatop=np.array([0.01, 0.010, 0.008, 0.024, 0.012, 0.012, 0.030, 0.030, 0.030, 0.012, 0.030, 0.030,
                         0.030, 0.030, 0.030, 0.030, 0.030]),
        leafcn=np.array([25.0, 58.0, 58.0, 25.8, 29.6, 29.6, 23.1, 23.1, 23.1, 29.6, 23.1, 23.1, 28.1, 28.1,
                         35.4, 28.0, 28.0]),
        flnr=np.array([0.1, 0.0509, 0.0466, 0.0546, 0.0461, 0.0515, 0.0716, 0.1007, 0.1007, 0.0517, 0.0943,
                       0.0943, 0.1365, 0.1365, 0.0900, 0.1200, 0.1200]),
        fnitr=np.full(n, 1.0),
        mbbopt=np.where(np.arange(n) == 14, 4.0, 9.0),
        medlynslope=np.array([4.0, 2.35, 2.35, 2.35, 4.12, 4.12, 4.45, 4.45, 4.45, 4.70, 4.70, 4.70, 2.22,
                              5.25, 1.62, 5.79, 5.79]),
        medlynintercept=np.full(n, 100.0),
        smpso=np.array([-66000.0, -66000.0, -66000.0, -66000.0, -66000.0, -35000.0, -35000.0, -35000.0,
                        -35000.0, -83000.0, -83000.0, -83000.0, -74000.0, -74000.0, -74000.0, -74000.0,
                        -74000.0]),
        smpsc=np.array([-255000.0, -255000.0, -255000.0, -255000.0, -255000.0, -224000.0, -224000.0,
                        -224000.0, -224000.0, -428000.0, -428000.0, -428000.0, -275000.0, -275000.0,
                        -275000.0, -275000.0, -275000.0]),
        theta_cj=np.where(np.arange(n) == 14, 0.80, 0.98),
        z0v_Cr=np.full(n, 0.35), z0v_Cs=np.full(n, 0.003), z0v_c=np.full(n, 0.37), z0v_cw=np.full(n, 2.0),
        z0v_LAImax=np.full(n, 2.0),
        is_tree=trees, is_shrub=shrubs,
        dbh=np.where(trees == 1, 0.35, np.where(shrubs == 1, 0.06, 0.0)),
        fbw=np.where(trees + shrubs > 0, 0.65, 0.7),
        nstem=np.where(trees == 1, 0.1, np.where(shrubs == 1, 1.0, 0.0)),
        rstem_per_dbh=np.full(n, 5.0e3),
        wood_density=np.where(trees + shrubs > 0, 500.0, 0.0),
    )
    return t


def _qsat_np(T, p):
    """Flatau et al. (1992) saturation humidity (input generation only)."""
    a = [6.11213476, 0.444007856, 0.143064234e-01, 0.264461437e-03, 0.305903558e-05, 0.196237241e-07,
         0.892344772e-10, -0.373208410e-12, 0.209339997e-15]
    c = [6.11123516, 0.503109514, 0.188369801e-01, 0.420547422e-03, 0.614396778e-05, 0.602780717e-07,
         0.387940929e-09, 0.149436277e-11, 0.262655803e-14]
    td = np.clip(T - TFRZ, -75.0, 100.0)
    k = np.where(td[None, :] >= 0.0, np.array(a)[:, None], np.array(c)[:, None])
    es = k[8]
    for i in range(7, -1, -1):
        es = k[i] + td * es
    es = es * 100.0
    return 0.622 * es / (p - 0.378 * es), es


def with_canopy(soil, seed, pmax=16, pmin=1, col0=0, gids=None):
    """Expand a soil-only state to 1..pmax vegetated patches per column with the
    canopy forcing and state of the full-column workloads (one gridcell per
    column).  Patch draws are keyed by global patch id col*pmax + k so any
    column regenerates alone."""
    d0 = soil.dims
    ncol = d0.ncol
    gid = _gids(ncol, col0, gids)
    npc = pmin + np.floor(uniform(seed, S_NPATCH, gid) * (pmax - pmin + 1)).astype(np.int32)
    npc = np.minimum(npc, pmax)
    npatch = int(npc.sum())
    d = Dims(ncol, npatch, d0.nlevsno, d0.nlevsoi, d0.nlevgrnd, ngrid=ncol)
    st = HostState(d)
    for k, v in soil.col.items():
        st.col[k] = v.copy()
    c, p, g = st.col, st.pch, st.can
    NS = d.nlevsno
    start = np.concatenate([[0], np.cumsum(npc)[:-1]]).astype(np.int64)
    c["patchi"][:] = (start + 1).astype(np.int32)
    c["patchf"][:] = (start + npc).astype(np.int32)
    pcol = np.repeat(np.arange(ncol), npc)                 # 0-based column of each patch
    kin = np.arange(npatch) - np.repeat(start, npc)       # patch rank within its column
    pid = gid[pcol] * pmax + kin                          # global patch id
    for k in g:
        g[k][...] = 0.0 if g[k].dtype == np.float64 else 0

    # ---- gridcell / column forcing
    ft = 250.0 + 65.0 * uniform(seed, S_FT, gid)
    pbot = 1000.0 * (60.0 + 43.0 * uniform(seed, S_PBOT, gid))
    rh = 20.0 + 80.0 * uniform(seed, S_RH, gid)
    qs, es = _qsat_np(ft, pbot)
    q = rh / 100.0 * qs
    wind = 0.5 + 14.5 * uniform(seed, S_WIND, gid)
    wdir = 2.0 * np.pi * uniform(seed, S_WDIR, gid)
    sw = 1000.0 * uniform(seed, S_SW, gid)
    lw = 150.0 + 300.0 * uniform(seed, S_LW, gid)
    th = ft * (1.0e5 / pbot) ** (RAIR / CPAIR)
    e = q * pbot / (0.622 + 0.378 * q)
    g["forc_u"][:] = wind * np.cos(wdir)
    g["forc_v"][:] = wind * np.sin(wdir)
    g["forc_pco2"][:] = 367.0e-6 * pbot
    g["forc_po2"][:] = 0.209 * pbot
    for k in ("forc_hgt_u", "forc_hgt_t", "forc_hgt_q"):
        g[k][:] = 30.0
    dayl = 3600.0 * (8.0 + 8.0 * uniform(seed, S_DAYL, gid))
    g["dayl"][:] = dayl
    g["max_dayl"][:] = 16.0 * 3600.0
    g["near_local_noon"][:] = (uniform(seed, S_NOON, gid) < 0.15).astype(np.int32)
    g["forc_q"][:] = q
    g["forc_pbot"][:] = pbot
    g["forc_th"][:] = th
    g["forc_rho"][:] = (pbot - 0.378 * e) / (RAIR * ft)
    g["thv"][:] = th * (1.0 + 0.61 * q)
    c["forc_lwrad"][:] = lw
    tg = c["t_grnd"]
    t1 = c["t_soisno"][NS]
    ttop = c["t_soisno"][NS + c["snl"], np.arange(ncol)]
    qsoil, _ = _qsat_np(t1, pbot)
    qsnow, _ = _qsat_np(ttop, pbot)
    qh2o, _ = _qsat_np(c["t_h2osfc"], pbot)
    beta = 0.2 + 0.8 * uniform(seed, S_BETA, gid)
    g["qg_soil"][:] = qsoil * beta
    # snow-free columns take the soil's humidity for the snow surface, so the
    # heat solve's top-layer flux equals the soil one (SurfaceHumidityMod.F90:198-203)
    g["qg_snow"][:] = np.where(c["snl"] < 0, qsnow, g["qg_soil"])
    g["qg_h2osfc"][:] = qh2o
    fs, fh = c["frac_sno_eff"], c["frac_h2osfc"]
    g["qg"][:] = fs * qsnow + (1.0 - fs - fh) * g["qg_soil"] + fh * qh2o
    qs1, _ = _qsat_np(tg + 0.5, pbot)
    qs0, _ = _qsat_np(tg - 0.5, pbot)
    g["dqgdT"][:] = (qs1 - qs0) * beta
    g["z0mg"][:] = 0.01
    g["soilbeta"][:] = beta
    g["soilresis"][:] = 50.0 + 1950.0 * uniform(seed, S_RESIS, gid)
    for k in ("h2osoi_liqvol", "qflx_tran_veg_col", "rootr_col"):
        g[k][...] = 0.0
    c["qflx_rootsoi"][...] = 0.0

    # ---- patches
    p["column"][:] = (pcol + 1).astype(np.int32)
    p["wtcol"][:] = 1.0 / npc[pcol]
    p["frac_veg_nosno"][:] = 1
    g["itype"][:] = (1 + np.floor(uniform(seed, S_PFT, pid) * 14)).astype(np.int32)
    g["gridcell"][:] = p["column"]
    g["active"][:] = 1
    elai = 8.0 * uniform(seed, S_LAI, pid)
    esai = 2.0 * uniform(seed, S_SAI, pid)
    elai = np.maximum(elai, 1.0e-3)
    htop = 0.1 + 34.9 * uniform(seed, S_HTOP, pid)
    kb = 0.5
    fsun = (1.0 - np.exp(-kb * elai)) / (kb * elai)
    g["elai"][:] = elai
    g["esai"][:] = esai
    g["laisun"][:] = elai * fsun
    g["laisha"][:] = elai - elai * fsun
    g["laisun_z"][:] = g["laisun"]
    g["laisha_z"][:] = g["laisha"]
    g["htop"][:] = htop
    swp = sw[pcol]
    lt = elai + esai
    g["sabv"][:] = 0.85 * swp * (1.0 - np.exp(-0.5 * lt))
    par = 0.5 * swp
    g["parsun_z"][:] = 0.5 * par * kb
    g["parsha_z"][:] = 0.1 * par * kb
    g["vcmaxcintsun"][:] = 0.6 + 0.4 * uniform(seed, S_VCS, pid)
    g["vcmaxcintsha"][:] = 0.3 + 0.4 * uniform(seed, S_VCH, pid)
    g["nrad"][:] = 1
    fhgt = 30.0 + 0.055 * htop + 0.67 * htop
    g["thm"][:] = ft[pcol] + 0.0098 * fhgt
    g["emv"][:] = 1.0 - np.exp(-lt / 1.0)
    fwet = 0.3 * uniform(seed, S_FWET, pid)
    g["fwet"][:] = fwet
    g["fdry"][:] = (1.0 - fwet) * elai / lt
    g["t_a10"][:] = ft[pcol] + 6.0 * (uniform(seed, S_T10, pid) - 0.5)
    for k in ("o3coefvsun", "o3coefvsha", "o3coefgsun", "o3coefgsha"):
        g[k][:] = 1.0
    # rooting profile (Zeng 2001 form) over the column's soil layers above bedrock
    zi = c["zi"][NS:, :]
    nbed = c["nbedrock"]
    ra = 6.0 + (g["itype"] % 5).astype(np.float64)
    rb = 1.5 + 0.1 * (g["itype"] % 7).astype(np.float64)
    rf = np.zeros((d.nlevgrnd, npatch))
    for j in range(1, d.nlevsoi + 1):
        z0, z1 = zi[j - 1, pcol], zi[j, pcol]
        v = 0.5 * (np.exp(-ra * z0) + np.exp(-rb * z0) - np.exp(-ra * z1) - np.exp(-rb * z1))
        rf[j - 1] = np.where(j <= nbed[pcol], v, 0.0)
    rf /= rf.sum(axis=0, keepdims=True)
    g["rootfr"][...] = rf
    g["displa"][:] = 0.67 * htop
    g["z0mv"][:] = 0.055 * htop
    tv = ft[pcol] + 4.0 * (uniform(seed, S_TV, pid) - 0.5)
    g["t_veg"][:] = tv
    g["t_stem"][:] = tv
    g["snocan"][:] = 0.0
    g["liqcan"][:] = 0.1 * uniform(seed, S_LIQCAN, pid)
    # ground heat inputs of SoilTemperature not produced by CanopyFluxes
    sabg = 0.8 * swp * np.exp(-0.5 * lt)
    for k in ("sabg", "sabg_soil", "sabg_snow"):
        p[k][:] = sabg
    p["sabg_lyr"][...] = 0.0
    p["sabg_lyr"][NS + c["snl"][pcol], np.arange(npatch)] = sabg
    for k in ("dlrad", "cgrnd", "eflx_sh_grnd", "eflx_sh_snow", "eflx_sh_soil", "eflx_sh_h2osfc",
              "qflx_evap_soi", "qflx_ev_snow", "qflx_ev_soil", "qflx_ev_h2osfc", "eflx_gnet", "dgnetdT",
              "sabg_chk"):
        p[k][:] = 0.0
    # SoilFluxes: temperatures saved before the step, outputs cleared
    for k in st.sf:
        st.sf[k][...] = 0.0
    st.sf["t_ssbef"][...] = c["t_soisno"]
    st.sf["t_h2osfc_bef"][:] = c["t_h2osfc"]
    with_hydrology(st, seed, col0=col0, gids=gids)
    return st


S_RAIN, S_ZWT, S_ZWTP, S_FROST, S_WTF, S_SLOPE, S_THRESH, S_EVAP, S_DEW = 70, 71, 72, 73, 74, 75, 76, 77, 78


def with_hydrology(st, seed, col0=0, gids=None):
    """Inputs of the surface water input chain (rain and snow melt, water
    tables, frost table, slope, micro-topography threshold, column
    evaporation fluxes), keyed by global column id; outputs cleared.  The
    column's frac_h2osfc / h2osfc are left as they are."""
    c, h = st.col, st.hy
    cols = _gids(st.dims.ncol, col0, gids)
    for k in h:
        h[k][...] = 0.0
    h["qflx_rain_plus_snomelt"][:] = 5e-4 * uniform(seed, S_RAIN, cols)
    zwt = 0.2 + 7.8 * uniform(seed, S_ZWT, cols)
    h["zwt"][:] = zwt
    h["zwt_perched"][:] = zwt * (0.2 + 0.8 * uniform(seed, S_ZWTP, cols))
    h["frost_table"][:] = 10.0 * uniform(seed, S_FROST, cols)
    h["wtfact"][:] = 0.1 + 0.4 * uniform(seed, S_WTF, cols)
    h["topo_slope"][:] = 30.0 * uniform(seed, S_SLOPE, cols)
    h["h2osfc_thresh"][:] = 5.0 + 45.0 * uniform(seed, S_THRESH, cols)
    h["frac_h2osfc_nosnow"][:] = c["frac_h2osfc"]
    for j, k in enumerate(("qflx_liqevap_from_top_layer", "qflx_ev_soil", "qflx_ev_h2osfc")):
        h[k][:] = -1e-6 + 5.1e-5 * uniform(seed, S_EVAP, cols, j)
    # dew, frost and sublimation of the top layer (RenewCondensation); the
    # sublimation never exceeds the ice there over an 1800 s step, as
    # SoilFluxes' limit guarantees in the model
    h["qflx_liqdew_to_top_layer"][:] = 1e-5 * uniform(seed, S_DEW, cols, 0)
    h["qflx_soliddew_to_top_layer"][:] = 1e-5 * uniform(seed, S_DEW, cols, 1)
    ice1 = c["h2osoi_ice"][st.dims.nlevsno]
    h["qflx_solidevap_from_top_layer"][:] = uniform(seed, S_DEW, cols, 2) * (
        h["qflx_soliddew_to_top_layer"] + ice1 / 1800.0)
    return st


CANOPY_SEEDS = {"canopy": 0xC75D0002, "canopy_uneven": 0xC75D0003}


def canopy_workload(name, ncol=None, seed=None, gids=None):
    """'canopy' (no snow, 20 layers, 1-16 patches) and 'canopy_uneven' (snow and
    bedrock mixes of the uneven workload, 1-16 patches).  gids: generate these
    global columns (in this order) instead of 0 .. ncol-1."""
    if gids is not None:
        ncol = len(gids)
    if name == "canopy":
        s = CANOPY_SEEDS[name] if seed is None else seed
        return with_canopy(soil_only(ncol or 262144, s, gids=gids), s, gids=gids)
    if name == "canopy_uneven":
        s = CANOPY_SEEDS[name] if seed is None else seed
        return with_canopy(soil_only(ncol or 1048576, s, snow=True, bedrock=True, gids=gids), s, gids=gids)
    raise KeyError(name)


def binned_ids(name, ncol, parts=4, seed=None):
    """Global ids 0 .. ncol-1 of a canopy workload in binned order: sorted by
    (snow layers, bedrock index) -- the work of a column's heat and Richards
    solves -- then dealt over `parts` contiguous parts (part k takes every
    parts-th column of the sorted order), so each part has the same mix."""
    s = CANOPY_SEEDS[name] if seed is None else seed
    uneven = name == "canopy_uneven"
    nsno, nbed = structure(ncol, s, snow=uneven, bedrock=uneven)
    o = np.lexsort((nbed, nsno))
    if parts > 1:
        o = np.concatenate([o[k::parts] for k in range(parts)])
    return o.astype(np.int64)
