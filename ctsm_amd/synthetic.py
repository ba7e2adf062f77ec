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
S_SAND, S_CLAY, S_OM, S_LIQ, S_T, S_INFL, S_SINKTOT, S_SINKW, S_G, S_NSNO, S_SDZ, S_SRHO, S_ST, S_BED = range(14)


def soil_only(ncol, seed, snow=False, bedrock=False, col0=0, nlevsno=12, nlevsoi=20, nlevgrnd=25,
              frac_h2osfc=None):
    """Columns of the soil-only workloads (one patch per column).

    col0 is the global id of the first column (for regenerating a slice)."""
    d = Dims(ncol, ncol, nlevsno, nlevsoi, nlevgrnd)
    st = HostState(d)
    c, p = st.col, st.pch
    gid = np.arange(col0, col0 + ncol, dtype=np.int64)
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
