// capi.hip -- the extern "C" boundary of libctsm_mi355x.so (include/ctsm_gpu.h).
//
// A context owns a HIP stream, the device error word, the heat-solve
// workspace and, for the drop-in (host-array) entry points, device mirrors
// of the fields a call touches.  Drop-in calls copy the filtered column range
// of every field to the device (strided 2-D copies over the Fortran arrays),
// launch, and copy the writable fields back before returning, which keeps the
// reference's synchronous `intent(inout)` semantics at the call site.
#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <vector>

#include "common.h"

namespace ctsm {
void launch_soilwater(hipStream_t, const View &, const Params &, const ctsm_col_state &, int, const int32_t *, double, ErrWord *,
                      int lanes, int block);
void launch_soiltemp_fractions(hipStream_t, const View &, const Params &, const ctsm_col_state &,
                               const ctsm_patch_state &, int, const int32_t *, double, double *, long, ErrWord *,
                               int block);
void launch_soilwater_fractions(hipStream_t, const View &, const ctsm_col_state &, int, const int32_t *);
void launch_soiltemp(hipStream_t, const View &, const Params &, const ctsm_col_state &, const ctsm_patch_state &, int,
                     const int32_t *, double, double *, long, ErrWord *);
void launch_soil_step(hipStream_t, const View &, const Params &, const ctsm_col_state &, const ctsm_patch_state &, int,
                      const int32_t *, double, double *, long, ErrWord *);
void launch_soiltemp_plain(hipStream_t, const View &, const Params &, const ctsm_col_state &, const ctsm_patch_state &,
                           int, const int32_t *, double, double *, long, ErrWord *);
size_t soiltemp_workspace_doubles(int n);
size_t richards_order_bytes(int n);
const int32_t *launch_richards_order(hipStream_t, const View &, const ctsm_col_state &, int, const int32_t *, void *,
                                     int);
size_t soiltemp_plain_workspace_doubles(int n);
size_t soilwater_plain_workspace_doubles(int n);
void launch_soilwater_plain(hipStream_t, const View &, const Params &, const ctsm_col_state &, int, const int32_t *,
                            double, double *, long, ErrWord *);
size_t canopy_work_bytes(int n, int itmax, int first_its, int round_its);
void launch_canopy(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_canopy_params &, const double *,
                   const ctsm_col_state &, const ctsm_canopy_state &, int, const int32_t *, double, ErrWord *,
                   void *, int, int, bool, int, bool, const ctsm_phs_state *, const ctsm_phs_canopy_state *,
                   const double *);
void launch_root_sink(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_col_state &,
                      const ctsm_canopy_state &, int, const int32_t *, ErrWord *);
void launch_save_before(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_col_state &,
                        const ctsm_soilflux_state &, int, const int32_t *);
void launch_soilfluxes(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_col_state &,
                       const ctsm_patch_state &, const ctsm_soilflux_state &, int, const int32_t *, int,
                       const int32_t *, double, ErrWord *);
void launch_balance(hipStream_t, const ctsm_layout &, const ctsm_config &, const double *, int, const int32_t *,
                    unsigned long long *);
void launch_energy_balance(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_col_state &,
                           const ctsm_patch_state &, const ctsm_canopy_state &, const ctsm_soilflux_state &, int,
                           const int32_t *, double *, double *, unsigned long long *, ErrWord *);
void launch_water_mass(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_col_state &,
                       const ctsm_canopy_state &, const ctsm_water_balance_state &, int, const int32_t *, double, int,
                       double *, ErrWord *);
void launch_water_balance(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_water_balance_state &, int,
                          const int32_t *, double, unsigned long long *, ErrWord *);
void launch_solar_balance(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_canopy_state &, int,
                          const int32_t *, const double *, const double *, const double *, const double *, double *,
                          unsigned long long *, ErrWord *);
void launch_surface_water(hipStream_t, const ctsm_layout &, double, const ctsm_hydro_params &, const ctsm_col_state &,
                          const ctsm_surface_water_state &, int, const int32_t *, double, ErrWord *);
void launch_patch2col3(hipStream_t, const ctsm_layout &, const ctsm_col_state &, const int32_t *, const double *,
                       const double *, const double *, const double *, double *, double *, double *, int,
                       const int32_t *, ErrWord *);
void launch_water_table(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_col_state &,
                        const ctsm_water_table_state &, int, const int32_t *, double, ErrWord *);
void launch_hist(hipStream_t, const ctsm_hist_field *, int, int);
void launch_drainage(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_hydro_params &,
                     const ctsm_col_state &, const ctsm_drainage_state &, int, const int32_t *, double, ErrWord *);
void launch_phs_conductance(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_phs_state &,
                            const double *, int, int, const int32_t *, ErrWord *);
void launch_phs_root_sink(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_phs_state &, int,
                          const int32_t *, ErrWord *);
void launch_phs_calcstress(hipStream_t, const ctsm_layout &, const ctsm_config &, const ctsm_phs_stress_state &,
                           const double *, int, int, const int32_t *, ErrWord *);
int ckpt_write(hipStream_t, const char *, const ctsm_ckpt_field *, int, hipError_t *);
int ckpt_read(hipStream_t, const char *, const ctsm_ckpt_field *, int, hipError_t *);
}  // namespace ctsm

using namespace ctsm;

namespace {

enum Kind { K_C, K_SNO, K_ZI, K_SOI, K_SNOW, K_P, K_LYR, K_G, K_SOIP, K_VEG, K_PHS2 };

struct FieldMeta {
    size_t offset;   // offset of the pointer member in the state struct
    Kind kind;
    int elem;        // bytes per element
    bool writable;   // copied back after a drop-in call
};

#define CF(name, kind, elem, wr) {offsetof(ctsm_col_state, name), kind, elem, wr}
const FieldMeta kColFields[] = {
    CF(snl, K_C, 4, false), CF(nbedrock, K_C, 4, false), CF(patchi, K_C, 4, false), CF(patchf, K_C, 4, false),
    CF(dz, K_SNO, 8, false), CF(z, K_SNO, 8, false), CF(zi, K_ZI, 8, false),
    CF(watsat, K_SOI, 8, false), CF(bsw, K_SOI, 8, false), CF(sucsat, K_SOI, 8, false), CF(hksat, K_SOI, 8, false),
    CF(tkmg, K_SOI, 8, false), CF(tkdry, K_SOI, 8, false), CF(csol, K_SOI, 8, false),
    CF(eff_porosity, K_SOI, 8, true), CF(smp_l, K_SOI, 8, true), CF(hk_l, K_SOI, 8, true), CF(thk, K_SNO, 8, true),
    CF(icefrac, K_SOI, 8, true), CF(qcharge, K_C, 8, true), CF(num_substeps, K_C, 8, true),
    CF(qin, K_SOI, 8, true), CF(qout, K_SOI, 8, true),
    CF(h2osoi_liq, K_SNO, 8, true), CF(h2osoi_ice, K_SNO, 8, true), CF(h2osfc, K_C, 8, true),
    CF(h2osno_no_layers, K_C, 8, true), CF(int_snow, K_C, 8, true),
    CF(frac_sno_eff, K_C, 8, false), CF(frac_h2osfc, K_C, 8, false), CF(snow_depth, K_C, 8, true),
    CF(bw, K_SNOW, 8, true), CF(snomelt_accum, K_C, 8, true),
    CF(qflx_infl, K_C, 8, false), CF(qflx_rootsoi, K_SOI, 8, false), CF(qflx_snomelt, K_C, 8, true),
    CF(qflx_snofrz, K_C, 8, true), CF(qflx_snow_drain, K_C, 8, true), CF(qflx_snomelt_lyr, K_SNOW, 8, true),
    CF(qflx_snofrz_lyr, K_SNOW, 8, true), CF(qflx_h2osfc_to_ice, K_C, 8, true),
    CF(t_soisno, K_SNO, 8, true), CF(t_h2osfc, K_C, 8, true), CF(t_grnd, K_C, 8, true), CF(fact, K_SNO, 8, true),
    CF(xmf, K_C, 8, true), CF(xmf_h2osfc, K_C, 8, true), CF(c_h2osfc, K_C, 8, true), CF(imelt, K_SNO, 4, true),
    CF(emg, K_C, 8, false), CF(htvp, K_C, 8, false), CF(eflx_bot, K_C, 8, false), CF(eflx_fgr12, K_C, 8, true),
    CF(eflx_fgr, K_SOI, 8, true), CF(eflx_snomelt, K_C, 8, true), CF(eflx_h2osfc_to_snow, K_C, 8, true),
    CF(forc_lwrad, K_C, 8, false), CF(zwt, K_C, 8, false), CF(smpmin, K_C, 8, false),
};
#undef CF
#define PF(name, kind, elem, wr) {offsetof(ctsm_patch_state, name), kind, elem, wr}
const FieldMeta kPatchFields[] = {
    PF(column, K_P, 4, false), PF(wtcol, K_P, 8, false), PF(frac_veg_nosno, K_P, 4, false),
    PF(sabg, K_P, 8, false), PF(sabg_soil, K_P, 8, false), PF(sabg_snow, K_P, 8, false), PF(sabg_lyr, K_LYR, 8, false),
    PF(dlrad, K_P, 8, false), PF(cgrnd, K_P, 8, false), PF(eflx_sh_grnd, K_P, 8, false),
    PF(eflx_sh_snow, K_P, 8, false), PF(eflx_sh_soil, K_P, 8, false), PF(eflx_sh_h2osfc, K_P, 8, false),
    PF(qflx_evap_soi, K_P, 8, false), PF(qflx_ev_snow, K_P, 8, false), PF(qflx_ev_soil, K_P, 8, false),
    PF(qflx_ev_h2osfc, K_P, 8, false), PF(eflx_gnet, K_P, 8, true), PF(dgnetdT, K_P, 8, true),
    PF(sabg_chk, K_P, 8, true),
};
#undef PF
// CanopyFluxes state, split by the index space each member lives in
#define CN(name, kind, elem, wr) {offsetof(ctsm_canopy_state, name), kind, elem, wr}
const FieldMeta kCanopyPatch[] = {
    CN(itype, K_P, 4, false), CN(column, K_P, 4, false), CN(gridcell, K_P, 4, false),
    CN(wtcol, K_P, 8, false), CN(frac_veg_nosno, K_P, 4, false), CN(active, K_P, 4, false),
    CN(elai, K_P, 8, false), CN(esai, K_P, 8, false), CN(laisun, K_P, 8, false),
    CN(laisha, K_P, 8, false), CN(htop, K_P, 8, false), CN(sabv, K_P, 8, false),
    CN(parsun_z, K_P, 8, false), CN(parsha_z, K_P, 8, false), CN(laisun_z, K_P, 8, false),
    CN(laisha_z, K_P, 8, false), CN(vcmaxcintsun, K_P, 8, false), CN(vcmaxcintsha, K_P, 8, false),
    CN(nrad, K_P, 4, false), CN(thm, K_P, 8, false), CN(emv, K_P, 8, false),
    CN(fwet, K_P, 8, false), CN(fdry, K_P, 8, false), CN(t_a10, K_P, 8, false),
    CN(o3coefvsun, K_P, 8, false), CN(o3coefvsha, K_P, 8, false), CN(o3coefgsun, K_P, 8, false),
    CN(o3coefgsha, K_P, 8, false), CN(rootfr, K_SOIP, 8, false), CN(displa, K_P, 8, true),
    CN(z0mv, K_P, 8, true), CN(t_veg, K_P, 8, true), CN(t_stem, K_P, 8, true),
    CN(snocan, K_P, 8, true), CN(liqcan, K_P, 8, true), CN(cgrnds, K_P, 8, true),
    CN(cgrndl, K_P, 8, true), CN(z0hv, K_P, 8, true), CN(z0qv, K_P, 8, true),
    CN(forc_hgt_u_patch, K_P, 8, true), CN(forc_hgt_t_patch, K_P, 8, true), CN(forc_hgt_q_patch, K_P, 8, true),
    CN(ram1, K_P, 8, true), CN(rb1, K_P, 8, true), CN(rah1, K_P, 8, true),
    CN(rah2, K_P, 8, true), CN(raw1, K_P, 8, true), CN(raw2, K_P, 8, true),
    CN(ustar, K_P, 8, true), CN(um, K_P, 8, true), CN(uaf, K_P, 8, true),
    CN(taf, K_P, 8, true), CN(qaf, K_P, 8, true), CN(obu, K_P, 8, true),
    CN(zeta, K_P, 8, true), CN(vpd, K_P, 8, true), CN(num_iter, K_P, 8, true),
    CN(u10, K_P, 8, true), CN(u10_clm, K_P, 8, true), CN(va, K_P, 8, true),
    CN(fv, K_P, 8, true), CN(vds, K_P, 8, true), CN(btran, K_P, 8, true),
    CN(rresis, K_SOIP, 8, true), CN(rootr, K_SOIP, 8, true), CN(eflx_sh_veg, K_P, 8, true),
    CN(eflx_sh_stem, K_P, 8, true), CN(eflx_sh_grnd, K_P, 8, true), CN(eflx_sh_snow, K_P, 8, true),
    CN(eflx_sh_soil, K_P, 8, true), CN(eflx_sh_h2osfc, K_P, 8, true), CN(taux, K_P, 8, true),
    CN(tauy, K_P, 8, true), CN(dlrad, K_P, 8, true), CN(ulrad, K_P, 8, true),
    CN(cgrnd, K_P, 8, true), CN(dhsdt_canopy, K_P, 8, true), CN(qflx_evap_soi, K_P, 8, true),
    CN(qflx_ev_snow, K_P, 8, true), CN(qflx_ev_soil, K_P, 8, true), CN(qflx_ev_h2osfc, K_P, 8, true),
    CN(qflx_tran_veg, K_P, 8, true), CN(qflx_evap_veg, K_P, 8, true), CN(rh_af, K_P, 8, true),
    CN(q_ref2m, K_P, 8, true), CN(rh_ref2m, K_P, 8, true), CN(rh_ref2m_r, K_P, 8, true),
    CN(vpd_ref2m, K_P, 8, true), CN(iwue_ln, K_P, 8, true), CN(t_ref2m, K_P, 8, true),
    CN(t_ref2m_r, K_P, 8, true), CN(t_skin, K_P, 8, true), CN(stem_biomass, K_P, 8, true),
    CN(leaf_biomass, K_P, 8, true), CN(dleaf_patch, K_P, 8, true), CN(rssun, K_P, 8, true),
    CN(rssha, K_P, 8, true), CN(psnsun, K_P, 8, true), CN(psnsha, K_P, 8, true),
    CN(psnsun_wc, K_P, 8, true), CN(psnsun_wj, K_P, 8, true), CN(psnsun_wp, K_P, 8, true),
    CN(psnsha_wc, K_P, 8, true), CN(psnsha_wj, K_P, 8, true), CN(psnsha_wp, K_P, 8, true),
    CN(fpsn, K_P, 8, true), CN(fpsn_wc, K_P, 8, true), CN(fpsn_wj, K_P, 8, true),
    CN(fpsn_wp, K_P, 8, true), CN(lmrsun, K_P, 8, true), CN(lmrsha, K_P, 8, true),
    CN(gs_mol_sun, K_P, 8, true), CN(gs_mol_sha, K_P, 8, true), CN(gs_mol_sun_ln, K_P, 8, true),
    CN(gs_mol_sha_ln, K_P, 8, true), CN(cisun_z, K_P, 8, true), CN(cisha_z, K_P, 8, true),
    CN(rssun_z, K_P, 8, true), CN(rssha_z, K_P, 8, true), CN(lmrsun_z, K_P, 8, true),
    CN(lmrsha_z, K_P, 8, true), CN(psnsun_z, K_P, 8, true), CN(psnsha_z, K_P, 8, true),
    CN(vcmax_z, K_P, 8, true), CN(tpu_z, K_P, 8, true), CN(kp_z, K_P, 8, true),
    CN(ac, K_P, 8, true), CN(aj, K_P, 8, true), CN(ap, K_P, 8, true),
    CN(ag, K_P, 8, true), CN(an, K_P, 8, true), CN(gb_mol, K_P, 8, true),
    CN(gs_mol, K_P, 8, true), CN(cp, K_P, 8, true), CN(kc, K_P, 8, true),
    CN(ko, K_P, 8, true), CN(qe, K_P, 8, true), CN(bbb, K_P, 8, true),
    CN(mbb, K_P, 8, true), CN(rh_leaf, K_P, 8, true), CN(vpd_can, K_P, 8, true),
    CN(lnca, K_P, 8, true), CN(c3flag, K_P, 4, true), CN(canopy_err, K_P, 8, true),
};
const FieldMeta kCanopyCol[] = {
    CN(forc_q, K_C, 8, false), CN(forc_pbot, K_C, 8, false), CN(forc_th, K_C, 8, false),
    CN(forc_rho, K_C, 8, false), CN(thv, K_C, 8, false), CN(qg, K_C, 8, false),
    CN(qg_snow, K_C, 8, false), CN(qg_soil, K_C, 8, false), CN(qg_h2osfc, K_C, 8, false),
    CN(dqgdT, K_C, 8, false), CN(z0mg, K_C, 8, false), CN(soilbeta, K_C, 8, false),
    CN(soilresis, K_C, 8, false), CN(h2osoi_liqvol, K_SOI, 8, true), CN(qflx_tran_veg_col, K_C, 8, true),
    CN(rootr_col, K_SOI, 8, true), CN(qflx_rootsoi_col, K_SOI, 8, true),
};
const FieldMeta kCanopyGrid[] = {
    CN(forc_u, K_G, 8, false), CN(forc_v, K_G, 8, false), CN(forc_pco2, K_G, 8, false),
    CN(forc_po2, K_G, 8, false), CN(forc_hgt_u, K_G, 8, false), CN(forc_hgt_t, K_G, 8, false),
    CN(forc_hgt_q, K_G, 8, false), CN(dayl, K_G, 8, false), CN(max_dayl, K_G, 8, false),
    CN(near_local_noon, K_G, 4, false),
};
#undef CN
// SoilFluxes state, split by index space
#define SFP(name, elem, wr) {offsetof(ctsm_soilflux_state, name), K_P, elem, wr}
const FieldMeta kSfPatch[] = {
    SFP(active, 4, false),
    SFP(eflx_sh_grnd, 8, true), SFP(qflx_evap_soi, 8, true), SFP(qflx_ev_snow, 8, true), SFP(qflx_ev_soil, 8, true),
    SFP(qflx_ev_h2osfc, 8, true), SFP(cgrnds, 8, false), SFP(cgrndl, 8, false), SFP(ulrad, 8, false),
    SFP(eflx_sh_veg, 8, false), SFP(eflx_sh_stem, 8, false), SFP(qflx_evap_veg, 8, false), SFP(qflx_tran_veg, 8, false),
    SFP(eflx_soil_grnd, 8, true), SFP(eflx_soil_grnd_r, 8, true), SFP(eflx_sh_tot, 8, true),
    SFP(eflx_sh_tot_r, 8, true), SFP(qflx_evap_tot, 8, true), SFP(eflx_lh_tot, 8, true), SFP(eflx_lh_tot_r, 8, true),
    SFP(qflx_evap_can, 8, true), SFP(eflx_lh_vege, 8, true), SFP(eflx_lh_vegt, 8, true), SFP(eflx_lh_grnd, 8, true),
    SFP(eflx_lwrad_out, 8, true), SFP(eflx_lwrad_out_r, 8, true), SFP(eflx_lwrad_net, 8, true),
    SFP(eflx_lwrad_net_r, 8, true), SFP(t_skin, 8, true), SFP(qflx_liqevap_from_top_layer, 8, true),
    SFP(qflx_solidevap_from_top_layer, 8, true), SFP(qflx_liqdew_to_top_layer, 8, true),
    SFP(qflx_soliddew_to_top_layer, 8, true), SFP(errsoi_patch, 8, true),
};
#undef SFP
const FieldMeta kSfCol[] = {
    {offsetof(ctsm_soilflux_state, t_ssbef), K_SNO, 8, false},
    {offsetof(ctsm_soilflux_state, t_h2osfc_bef), K_C, 8, false},
    {offsetof(ctsm_soilflux_state, errsoi_col), K_C, 8, true},
};
// surface water input chain: column members only
#define SW(name, wr) {offsetof(ctsm_surface_water_state, name), K_C, 8, wr}
const FieldMeta kSw[] = {
    SW(qflx_rain_plus_snomelt, false), SW(qflx_snow_h2osfc, false), SW(qflx_floodc, false),
    SW(qflx_liqevap_from_top_layer, false), SW(qflx_ev_soil, false), SW(qflx_ev_h2osfc, false),
    SW(zwt, false), SW(zwt_perched, false), SW(frost_table, false), SW(wtfact, false),
    SW(topo_slope, false), SW(h2osfc_thresh, false), SW(frac_h2osfc_nosnow, false), SW(h2osfc, true),
    SW(fsat, true), SW(fcov, true), SW(qflx_sat_excess_surf, true), SW(qflx_top_soil, true), SW(qflx_in_soil, true),
    SW(qflx_top_soil_to_h2osfc, true), SW(qinmax, true), SW(qflx_infl_excess, true), SW(qflx_in_soil_limited, true),
    SW(qflx_in_h2osfc, true), SW(qflx_infl_excess_surf, true), SW(qflx_h2osfc_surf, true),
    SW(qflx_h2osfc_drain, true), SW(qflx_infl, true), SW(qflx_surf, true),
};
#undef SW
constexpr int kNSw = sizeof(kSw) / sizeof(kSw[0]);
static_assert(kNSw * sizeof(void *) == sizeof(ctsm_surface_water_state),
              "every ctsm_surface_water_state member has a mirror entry");
// water tables and condensation
const FieldMeta kWt[] = {
    {offsetof(ctsm_water_table_state, zwt), K_C, 8, true},
    {offsetof(ctsm_water_table_state, zwt_perched), K_C, 8, true},
    {offsetof(ctsm_water_table_state, frost_table), K_C, 8, true},
    {offsetof(ctsm_water_table_state, h2osoi_vol), K_SOI, 8, true},
    {offsetof(ctsm_water_table_state, qflx_liqdew_to_top_layer), K_C, 8, false},
    {offsetof(ctsm_water_table_state, qflx_soliddew_to_top_layer), K_C, 8, false},
    {offsetof(ctsm_water_table_state, qflx_solidevap_from_top_layer), K_C, 8, false},
};
constexpr int kNWt = sizeof(kWt) / sizeof(kWt[0]);
static_assert(kNWt * sizeof(void *) == sizeof(ctsm_water_table_state),
              "every ctsm_water_table_state member has a mirror entry");
// drainage
#define DR(name, kind, wr) {offsetof(ctsm_drainage_state, name), kind, 8, wr}
const FieldMeta kDr[] = {
    DR(topo_slope, K_C, false), DR(zwt_perched, K_C, false), DR(frost_table, K_C, false), DR(qflx_surf, K_C, false),
    DR(qflx_snwcp_liq, K_C, false), DR(zwt, K_C, true), DR(h2osoi_vol, K_SOI, true),
    DR(qflx_drain_perched, K_C, true), DR(qflx_latflow_out, K_C, true), DR(qflx_latflow_in, K_C, true),
    DR(qflx_rsub_sat, K_C, true), DR(qflx_ice_runoff_xs, K_C, true), DR(qflx_drain, K_C, true),
    DR(qflx_qrgwl, K_C, true), DR(qflx_runoff, K_C, true),
};
#undef DR
constexpr int kNDr = sizeof(kDr) / sizeof(kDr[0]);
static_assert(kNDr * sizeof(void *) == sizeof(ctsm_drainage_state),
              "every ctsm_drainage_state member has a mirror entry");
// plant hydraulic stress, patch members then column members
#define PH(name, kind, elem, wr) {offsetof(ctsm_phs_state, name), kind, elem, wr}
const FieldMeta kPhsP[] = {
    PH(itype, K_P, 4, false), PH(column, K_P, 4, false), PH(active, K_P, 4, false),
    PH(frac_veg_nosno, K_P, 4, false), PH(wtcol, K_P, 8, false), PH(froot_carbon, K_P, 8, false),
    PH(tsai, K_P, 8, false), PH(tlai, K_P, 8, false), PH(rootfr, K_SOIP, 8, false), PH(vegwp, K_VEG, 8, false),
    PH(k_soil_root, K_SOIP, 8, true), PH(root_conductance, K_SOIP, 8, true), PH(soil_conductance, K_SOIP, 8, true),
    PH(qflx_hydr_redist, K_P, 8, true),
};
const FieldMeta kPhsC[] = {
    PH(patchi, K_C, 4, false), PH(patchf, K_C, 4, false), PH(dz, K_SNO, 8, false), PH(z, K_SNO, 8, false),
    PH(hksat, K_SOI, 8, false), PH(hk_l, K_SOI, 8, false), PH(smp_l, K_SOI, 8, false),
    PH(qflx_rootsoi, K_SOI, 8, true), PH(qflx_phs_neg, K_C, 8, true),
};
#undef PH
constexpr int kNPhsP = sizeof(kPhsP) / sizeof(kPhsP[0]);
constexpr int kNPhsC = sizeof(kPhsC) / sizeof(kPhsC[0]);
static_assert((kNPhsP + kNPhsC) * sizeof(void *) == sizeof(ctsm_phs_state),
              "every ctsm_phs_state member has a mirror entry");
// calcstress: every member is used
#define PS(name, kind, elem, wr) {offsetof(ctsm_phs_stress_state, name), kind, elem, wr}
const FieldMeta kStP[] = {
    PS(itype, K_P, 4, false), PS(column, K_P, 4, false), PS(local_morning, K_P, 4, false),
    PS(laisun, K_P, 8, false), PS(laisha, K_P, 8, false), PS(elai, K_P, 8, false), PS(esai, K_P, 8, false),
    PS(tsai, K_P, 8, false), PS(htop, K_P, 8, false), PS(fdry, K_P, 8, false), PS(thm, K_P, 8, false),
    PS(gb_mol, K_P, 8, false), PS(gs_mol_sun, K_P, 8, false), PS(gs_mol_sha, K_P, 8, false),
    PS(qsatl, K_P, 8, false), PS(qaf, K_P, 8, false), PS(k_soil_root, K_SOIP, 8, false),
    PS(vegwp, K_VEG, 8, true), PS(vegwp_pd, K_VEG, 8, true), PS(bsun, K_P, 8, true), PS(bsha, K_P, 8, true),
    PS(qflx_tran_veg, K_P, 8, true),
};
const FieldMeta kStC[] = {
    PS(forc_rho, K_C, 8, false), PS(forc_pbot, K_C, 8, false), PS(z, K_SNO, 8, false), PS(smp_l, K_SOI, 8, false),
};
#undef PS
// CanopyFluxes with plant hydraulic stress: every ctsm_phs_canopy_state member
#define PC(name, kind, elem, wr) {offsetof(ctsm_phs_canopy_state, name), kind, elem, wr}
const FieldMeta kPhc[] = {
    PC(tlai_z, K_P, 8, false), PC(local_morning, K_P, 4, false), PC(vegwp, K_VEG, 8, true),
    PC(vegwp_ln, K_VEG, 8, true), PC(vegwp_pd, K_VEG, 8, true), PC(bsun, K_P, 8, true), PC(bsha, K_P, 8, true),
    PC(an_sun, K_P, 8, true), PC(an_sha, K_P, 8, true), PC(vcmax_z_phs, K_PHS2, 8, true),
    PC(tpu_z_phs, K_PHS2, 8, true), PC(kp_z_phs, K_PHS2, 8, true), PC(ac_phs, K_PHS2, 8, true),
    PC(aj_phs, K_PHS2, 8, true), PC(ap_phs, K_PHS2, 8, true), PC(ag_phs, K_PHS2, 8, true),
};
#undef PC
constexpr int kNPhc = sizeof(kPhc) / sizeof(kPhc[0]);
static_assert(kNPhc * sizeof(void *) == sizeof(ctsm_phs_canopy_state),
              "every ctsm_phs_canopy_state member has a mirror entry");
constexpr int kNStP = sizeof(kStP) / sizeof(kStP[0]);
constexpr int kNStC = sizeof(kStC) / sizeof(kStC[0]);
static_assert((kNStP + kNStC) * sizeof(void *) == sizeof(ctsm_phs_stress_state),
              "every ctsm_phs_stress_state member has a mirror entry");
constexpr int kNSfP = sizeof(kSfPatch) / sizeof(kSfPatch[0]);
constexpr int kNSfC = sizeof(kSfCol) / sizeof(kSfCol[0]);
static_assert((kNSfP + kNSfC) * sizeof(void *) == sizeof(ctsm_soilflux_state),
              "every ctsm_soilflux_state member has a mirror entry");
constexpr int kNCanP = sizeof(kCanopyPatch) / sizeof(kCanopyPatch[0]);
constexpr int kNCanC = sizeof(kCanopyCol) / sizeof(kCanopyCol[0]);
constexpr int kNCanG = sizeof(kCanopyGrid) / sizeof(kCanopyGrid[0]);
static_assert((kNCanP + kNCanC + kNCanG) * sizeof(void *) == sizeof(ctsm_canopy_state),
              "every ctsm_canopy_state member has a mirror entry");
static_assert(sizeof(kColFields) / sizeof(kColFields[0]) * sizeof(void *) == sizeof(ctsm_col_state),
              "every ctsm_col_state member has a mirror entry");

int nlev_of(Kind k, const ctsm_config &cfg)
{
    switch (k) {
    case K_C: case K_P: case K_G: return 1;
    case K_SOIP: return cfg.nlevgrnd;
    case K_VEG: return 4;
    case K_PHS2: return 2;
    case K_SNO: return cfg.nlevsno + cfg.nlevgrnd;
    case K_ZI: return cfg.nlevsno + cfg.nlevgrnd + 1;
    case K_SOI: return cfg.nlevgrnd;
    case K_SNOW: return cfg.nlevsno;
    case K_LYR: return cfg.nlevsno + 1;
    }
    return 0;
}

// Device buffer that only grows.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t b)
    {
        if (b <= bytes) return hipSuccess;
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, b);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    void release()
    {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

// Streaming copy: 16 B per lane per iteration, grid-stride.
__global__ void __launch_bounds__(256) k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n16)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
__global__ void k_copy1(char *__restrict__ dst, const char *__restrict__ src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Column indices begc .. begc + n - 1: the filter of "every column", so a
// step split into contiguous parts can address part k as out + k * len.
__global__ void __launch_bounds__(256) k_iota(int begc, int n, int32_t *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = begc + i;
}

__global__ void k_reset_err(ErrWord *e)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        e->first_col = INT32_MAX;
        e->code = 0;
    }
}

}  // namespace

// Scratch of ctsm_canopyfluxes_dev_noalloc (the compacted rounds' state,
// lists and counters).  In host code `scratch` is only a handle: kernels get
// the per-device address ctsm_create resolves with hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[CTSM_NOALLOC_SCRATCH_BYTES];

struct ctsm_ctx {
    int device = 0;
    void *scratch_dev = nullptr;      // this device's copy of `scratch`
    ctsm_config cfg{};
    Params prm{};
    ctsm_tuning tun{4, 128, 128, 0, 6, 3, 4, 0, 1};
    static constexpr int kMaxStreams = 4;
    hipStream_t side[kMaxStreams - 1] = {nullptr, nullptr, nullptr};   // concurrent parts of a step
    hipEvent_t ev_fork = nullptr, ev_join[kMaxStreams - 1] = {nullptr, nullptr, nullptr};
    DevBuf partf;                      // per-part column filters of a split step
    hipStream_t stream = nullptr;
    ErrWord *err = nullptr;            // device
    DevBuf work;                       // heat-solve workspace
    DevBuf filt;                       // filter (drop-in mode)
    std::vector<DevBuf> colmir, pchmir; // drop-in mirrors
    std::vector<DevBuf> canp, canc, cang; // canopy drop-in mirrors
    ctsm_canopy_params cprm{};         // canopy param-file scalars
    DevBuf pft;                        // PFT table (N_PFT_FIELDS x CTSM_MAX_PFT doubles)
    DevBuf cwork;                      // CanopyFluxes round state, patch lists and counters
    DevBuf filt2;                      // second filter (drop-in calls taking column and patch filters)
    DevBuf bal;                        // balance-check reduction word
    std::vector<DevBuf> sfp, sfc;      // SoilFluxes drop-in mirrors
    std::vector<DevBuf> swm;           // surface-water drop-in mirrors
    std::vector<DevBuf> wtm;           // water-table drop-in mirrors
    std::vector<DevBuf> drm;           // drainage drop-in mirrors
    std::vector<DevBuf> phsp, phsc;    // plant-hydraulics drop-in mirrors
    DevBuf phs_tab;                    // PHS parameter table (17 x CTSM_MAX_PFT doubles)
    int32_t phs_npft = 0;              // 0 until ctsm_set_phs_params
    bool phs_kmax = false;             // the table holds kmax (calcstress may run)
    std::vector<DevBuf> stp, stc;      // calcstress drop-in mirrors
    std::vector<DevBuf> phcm;          // PHS CanopyFluxes drop-in mirrors
    bool have_canopy = false;
    int32_t last_code = 0, last_index = 0;
    std::mutex mu;
};

namespace {

int hip_fail(ctsm_ctx *ctx, hipError_t e)
{
    if (ctx) { ctx->last_code = CTSM_E_HIP; ctx->last_index = (int32_t)e; }
    return CTSM_E_HIP;
}

Params make_params(const ctsm_config &c)
{
    Params p{};
    p.dtmin = c.dtmin;
    p.verySmall = c.verySmall;
    p.xTolerUpper = c.xTolerUpper;
    p.xTolerLower = c.xTolerLower;
    p.e_ice = c.e_ice;
    p.lower_bc = c.lower_boundary_condition;
    p.flux_calculation = c.flux_calculation;
    p.snow_thermal_cond = c.snow_thermal_cond_method;
    return p;
}

View make_view(const ctsm_ctx *ctx, const ctsm_layout &L)
{
    View v{};
    v.begc = L.begc_alloc;
    v.ldc = L.ldc;
    v.begp = L.begp_alloc;
    v.ldp = L.ldp;
    v.nlevsno = ctx->cfg.nlevsno;
    v.nlevsoi = ctx->cfg.nlevsoi;
    v.nlevgrnd = ctx->cfg.nlevgrnd;
    return v;
}

int check_layout(const ctsm_layout *L)
{
    if (!L || L->ldc <= 0 || L->ldp < 0) return CTSM_E_ARG;
    return CTSM_OK;
}

int ensure_work(ctsm_ctx *ctx, size_t doubles)
{
    hipError_t e = ctx->work.ensure(std::max<size_t>(doubles, 1) * sizeof(double));
    return e == hipSuccess ? CTSM_OK : hip_fail(ctx, e);
}

int finish_launch(ctsm_ctx *ctx)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? CTSM_OK : hip_fail(ctx, e);
}

int sync_and_check(ctsm_ctx *ctx)
{
    ErrWord h{};
    hipError_t e = hipMemcpyAsync(&h, ctx->err, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e);
    hipLaunchKernelGGL(k_reset_err, dim3(1), dim3(1), 0, ctx->stream, ctx->err);
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e);
    if (h.code != 0) {
        ctx->last_code = h.code;
        ctx->last_index = h.first_col;
        return h.code;
    }
    ctx->last_code = 0;
    ctx->last_index = 0;
    return CTSM_OK;
}

// ---- drop-in mirroring ------------------------------------------------------

struct Range { int lo, hi; };

Range filter_range(int n, const int32_t *f, int def_lo)
{
    if (!f) return {def_lo, def_lo + n - 1};
    int lo = INT_MAX, hi = INT_MIN;
    for (int i = 0; i < n; ++i) { lo = std::min(lo, (int)f[i]); hi = std::max(hi, (int)f[i]); }
    return {lo, hi};
}

// Copy [lo, hi] of every level of each non-null field between host arrays
// (layout ld / beg) and dense device mirrors (layout hi-lo+1 / lo).
int mirror(ctsm_ctx *ctx, const FieldMeta *meta, int nmeta, const void *host_struct, void *dev_struct,
           std::vector<DevBuf> &bufs, int beg_alloc, int ld, Range r, bool to_device)
{
    const int w = r.hi - r.lo + 1;
    if (bufs.size() < (size_t)nmeta) bufs.resize(nmeta);
    for (int f = 0; f < nmeta; ++f) {
        const FieldMeta &m = meta[f];
        char *hp = *(char *const *)((const char *)host_struct + m.offset);
        char **dslot = (char **)((char *)dev_struct + m.offset);
        if (!hp) { *dslot = nullptr; continue; }
        if (!to_device && !m.writable) continue;
        const int nl = nlev_of(m.kind, ctx->cfg);
        if (to_device) {
            hipError_t e = bufs[f].ensure((size_t)w * nl * m.elem);
            if (e != hipSuccess) return hip_fail(ctx, e);
            *dslot = (char *)bufs[f].p;
        }
        char *dp = (char *)bufs[f].p;
        const char *hsrc = hp + (size_t)(r.lo - beg_alloc) * m.elem;
        hipError_t e;
        if (to_device)
            e = hipMemcpy2DAsync(dp, (size_t)w * m.elem, hsrc, (size_t)ld * m.elem, (size_t)w * m.elem, nl,
                                 hipMemcpyHostToDevice, ctx->stream);
        else
            e = hipMemcpy2DAsync((void *)hsrc, (size_t)ld * m.elem, dp, (size_t)w * m.elem, (size_t)w * m.elem, nl,
                                 hipMemcpyDeviceToHost, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e);
    }
    return CTSM_OK;
}

// Patch range of the filtered columns (inclusive).
Range patch_range(const ctsm_layout *L, const ctsm_col_state *col, int n, const int32_t *f)
{
    int lo = INT_MAX, hi = INT_MIN;
    for (int i = 0; i < n; ++i) {
        const int c = f ? f[i] : L->begc_alloc + i;
        const int pi = col->patchi[c - L->begc_alloc], pf = col->patchf[c - L->begc_alloc];
        if (pf >= pi) { lo = std::min(lo, pi); hi = std::max(hi, pf); }
    }
    if (lo > hi) return {L->begp_alloc, L->begp_alloc};
    return {lo, hi};
}

enum Routine { R_SOILWATER, R_SOILTEMP, R_FRACTIONS };

int run_host(ctsm_ctx *ctx, Routine which, const ctsm_layout *L, const ctsm_col_state *col,
             const ctsm_patch_state *pch, int32_t n, const int32_t *filter, double dtime)
{
    if (check_layout(L) || !col || n < 0 || (which == R_SOILTEMP && !pch)) return CTSM_E_ARG;
    if (n == 0) return CTSM_OK;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    const Range cr = filter_range(n, filter, L->begc_alloc);
    if (cr.lo < L->begc_alloc || cr.hi >= L->begc_alloc + L->ldc) return CTSM_E_ARG;
    ctsm_col_state dcol{};
    ctsm_patch_state dpch{};
    int rc = mirror(ctx, kColFields, (int)(sizeof(kColFields) / sizeof(kColFields[0])), col, &dcol, ctx->colmir,
                    L->begc_alloc, L->ldc, cr, true);
    if (rc) return rc;
    Range pr{L->begp_alloc, L->begp_alloc};
    if (which == R_SOILTEMP) {
        pr = patch_range(L, col, n, filter);
        if (pr.lo < L->begp_alloc || pr.hi >= L->begp_alloc + L->ldp) return CTSM_E_ARG;
        rc = mirror(ctx, kPatchFields, (int)(sizeof(kPatchFields) / sizeof(kPatchFields[0])), pch, &dpch, ctx->pchmir,
                    L->begp_alloc, L->ldp, pr, true);
        if (rc) return rc;
    }
    const int32_t *dfilt = nullptr;
    if (filter) {
        he = ctx->filt.ensure((size_t)n * sizeof(int32_t));
        if (he != hipSuccess) return hip_fail(ctx, he);
        he = hipMemcpyAsync(ctx->filt.p, filter, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
        if (he != hipSuccess) return hip_fail(ctx, he);
        dfilt = (const int32_t *)ctx->filt.p;
    }
    ctsm_layout dl{cr.lo, cr.hi - cr.lo + 1, pr.lo, pr.hi - pr.lo + 1};
    View V = make_view(ctx, dl);
    switch (which) {
    case R_SOILWATER:
        launch_soilwater(ctx->stream, V, ctx->prm, dcol, n, dfilt, dtime, ctx->err, ctx->tun.richards_lanes,
                         ctx->tun.richards_block);
        break;
    case R_FRACTIONS:
        launch_soilwater_fractions(ctx->stream, V, dcol, n, dfilt);
        break;
    case R_SOILTEMP:
        if ((rc = ensure_work(ctx, soiltemp_workspace_doubles(n)))) return rc;
        launch_soiltemp(ctx->stream, V, ctx->prm, dcol, dpch, n, dfilt, dtime, (double *)ctx->work.p, n, ctx->err);
        break;
    }
    if ((rc = finish_launch(ctx))) return rc;
    rc = mirror(ctx, kColFields, (int)(sizeof(kColFields) / sizeof(kColFields[0])), col, &dcol, ctx->colmir,
                L->begc_alloc, L->ldc, cr, false);
    if (rc) return rc;
    if (which == R_SOILTEMP) {
        rc = mirror(ctx, kPatchFields, (int)(sizeof(kPatchFields) / sizeof(kPatchFields[0])), pch, &dpch, ctx->pchmir,
                    L->begp_alloc, L->ldp, pr, false);
        if (rc) return rc;
    }
    return sync_and_check(ctx);
}

}  // namespace

extern "C" {

void ctsm_config_default(ctsm_config *cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->nlevsno = 12;
    cfg->nlevsoi = 20;
    cfg->nlevgrnd = 25;
    cfg->soilwater_movement_method = 1;
    cfg->upper_boundary_condition = CTSM_BC_FLUX;
    cfg->lower_boundary_condition = CTSM_BC_ZERO_FLUX;
    cfg->flux_calculation = 1;
    cfg->snow_thermal_cond_method = CTSM_SNOWTHK_STURM1997;
    cfg->itmax_canopy_fluxes = 40;
    cfg->dtmin = 60.0;
    cfg->verySmall = 1.e-8;
    cfg->xTolerUpper = 1.e-1;
    cfg->xTolerLower = 1.e-2;
    cfg->e_ice = 6.0;
    cfg->z0param_method = CTSM_Z0_MEIER2022;
    cfg->stomatalcond_method = CTSM_STOMATA_MEDLYN2011;
    cfg->soil_resis_method = CTSM_SOILRESIS_SL14;
    cfg->use_biomass_heat_storage = 1;
    cfg->use_undercanopy_stability = 0;
    cfg->light_inhibit = 1;
    cfg->use_hydrstress = 0;
    cfg->use_excess_ice = 0;
    cfg->modifyphoto_and_lmr_forcrop = 1;
    cfg->pad1 = 0;
}

int ctsm_create(const ctsm_config *cfg, int device, ctsm_ctx **out)
{
    if (!cfg || !out) return CTSM_E_ARG;
    *out = nullptr;
    if (cfg->soilwater_movement_method != 1 || cfg->upper_boundary_condition != CTSM_BC_FLUX ||
        (cfg->lower_boundary_condition != CTSM_BC_FLUX && cfg->lower_boundary_condition != CTSM_BC_ZERO_FLUX &&
         cfg->lower_boundary_condition != CTSM_BC_WATERTABLE) ||
        cfg->flux_calculation != 1 ||
        (cfg->snow_thermal_cond_method != CTSM_SNOWTHK_JORDAN1991 &&
         cfg->snow_thermal_cond_method != CTSM_SNOWTHK_STURM1997) ||
        cfg->nlevsno < 1 || cfg->nlevsno > MAX_NLEVSNO || cfg->nlevsoi < 2 || cfg->nlevsoi > MAX_NLEVSOI ||
        cfg->nlevgrnd < cfg->nlevsoi || cfg->nlevgrnd > MAX_NLEVGRND || cfg->nlevsno + 1 + cfg->nlevgrnd > MAX_ROWS)
        return CTSM_E_CONFIG;
    if ((cfg->z0param_method != CTSM_Z0_ZENGWANG2007 && cfg->z0param_method != CTSM_Z0_MEIER2022) ||
        (cfg->stomatalcond_method != CTSM_STOMATA_BB1987 && cfg->stomatalcond_method != CTSM_STOMATA_MEDLYN2011) ||
        (cfg->soil_resis_method != CTSM_SOILRESIS_LEEPIELKE1992 && cfg->soil_resis_method != CTSM_SOILRESIS_SL14) ||
        (unsigned)cfg->use_biomass_heat_storage > 1u || (unsigned)cfg->use_undercanopy_stability > 1u ||
        (unsigned)cfg->light_inhibit > 1u || cfg->itmax_canopy_fluxes < 0 || cfg->itmax_canopy_fluxes > 1000)
        return CTSM_E_CONFIG;
    if ((unsigned)cfg->use_hydrstress > 1u || cfg->use_excess_ice != 0 ||
        (unsigned)cfg->modifyphoto_and_lmr_forcrop > 1u)
        return CTSM_E_CONFIG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CTSM_E_NODEVICE;
    if (device < 0 || device >= ndev) return CTSM_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return CTSM_E_HIP;
    ctsm_ctx *ctx = new ctsm_ctx();
    ctx->device = device;
    ctx->cfg = *cfg;
    ctx->prm = make_params(*cfg);
    bool ok = hipGetSymbolAddress(&ctx->scratch_dev, HIP_SYMBOL(scratch)) == hipSuccess && ctx->scratch_dev &&
              hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void **)&ctx->err, sizeof(ErrWord)) == hipSuccess &&
              hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < ctsm_ctx::kMaxStreams - 1; ++k)
        ok = hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&ctx->ev_join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        ctsm_destroy(ctx);
        return CTSM_E_HIP;
    }
    hipLaunchKernelGGL(k_reset_err, dim3(1), dim3(1), 0, ctx->stream, ctx->err);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
        ctsm_destroy(ctx);
        return CTSM_E_HIP;
    }
    *out = ctx;
    return CTSM_OK;
}

int ctsm_destroy(ctsm_ctx *ctx)
{
    if (!ctx) return CTSM_OK;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    ctx->work.release();
    ctx->filt.release();
    for (auto &b : ctx->colmir) b.release();
    for (auto &b : ctx->pchmir) b.release();
    for (auto &b : ctx->canp) b.release();
    for (auto &b : ctx->canc) b.release();
    for (auto &b : ctx->cang) b.release();
    ctx->pft.release();
    ctx->cwork.release();
    ctx->filt2.release();
    ctx->bal.release();
    for (auto &b : ctx->sfp) b.release();
    for (auto &b : ctx->sfc) b.release();
    for (auto &b : ctx->swm) b.release();
    for (auto &b : ctx->wtm) b.release();
    for (auto &b : ctx->drm) b.release();
    for (auto &b : ctx->phsp) b.release();
    for (auto &b : ctx->phsc) b.release();
    for (auto &b : ctx->stp) b.release();
    for (auto &b : ctx->stc) b.release();
    for (auto &b : ctx->phcm) b.release();
    ctx->phs_tab.release();
    ctx->partf.release();
    for (int k = 0; k < ctsm_ctx::kMaxStreams - 1; ++k) {
        if (ctx->side[k]) { (void)hipStreamSynchronize(ctx->side[k]); (void)hipStreamDestroy(ctx->side[k]); }
        if (ctx->ev_join[k]) (void)hipEventDestroy(ctx->ev_join[k]);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->err) hipFree(ctx->err);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
    return CTSM_OK;
}

int ctsm_set_tuning(ctsm_ctx *ctx, const ctsm_tuning *t)
{
    if (!ctx || !t) return CTSM_E_ARG;
    if (t->richards_lanes != 0 && t->richards_lanes != 1 && t->richards_lanes != 2 && t->richards_lanes != 4 &&
        t->richards_lanes != 8)
        return CTSM_E_ARG;
    if (t->richards_block < 0 || t->richards_block > 256 || (t->richards_block % 64) != 0) return CTSM_E_ARG;
    /* the one-lane Richards kernel is built for 64-lane blocks (register budget of one wave per SIMD) */
    if (t->richards_lanes == 1 && t->richards_block > 64) return CTSM_E_ARG;
    if (t->soiltemp_block < 0 || t->soiltemp_block > 128 || (t->soiltemp_block % 64) != 0) return CTSM_E_ARG;
    if (t->canopy_first_iterations < 1 || t->canopy_first_iterations > 1001 || t->canopy_round_iterations < 1 ||
        t->canopy_round_iterations > 1001)
        return CTSM_E_ARG;
    if (t->step_streams < 1 || t->step_streams > ctsm_ctx::kMaxStreams) return CTSM_E_ARG;
    if (t->plain_routines != 0 && t->plain_routines != 1) return CTSM_E_ARG;
    if (t->canopy_waves < 1 || t->canopy_waves > 3) return CTSM_E_ARG;
    ctx->tun = *t;
    return CTSM_OK;
}

int ctsm_get_tuning(const ctsm_ctx *ctx, ctsm_tuning *t)
{
    if (!ctx || !t) return CTSM_E_ARG;
    *t = ctx->tun;
    return CTSM_OK;
}

int ctsm_last_error(const ctsm_ctx *ctx, int32_t *code, int32_t *index)
{
    if (!ctx) return CTSM_E_ARG;
    if (code) *code = ctx->last_code;
    if (index) *index = ctx->last_index;
    return CTSM_OK;
}

int ctsm_soilwater(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col, int32_t n,
                   const int32_t *filter, double dtime)
{
    if (!ctx) return CTSM_E_ARG;
    return run_host(ctx, R_SOILWATER, lay, col, nullptr, n, filter, dtime);
}

int ctsm_soiltemperature(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                         const ctsm_patch_state *pch, int32_t n, const int32_t *filter, double dtime)
{
    if (!ctx) return CTSM_E_ARG;
    return run_host(ctx, R_SOILTEMP, lay, col, pch, n, filter, dtime);
}

int ctsm_soilwater_fractions(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col, int32_t n,
                             const int32_t *filter)
{
    if (!ctx) return CTSM_E_ARG;
    return run_host(ctx, R_FRACTIONS, lay, col, nullptr, n, filter, 0.0);
}

int ctsm_soilwater_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col, int32_t n,
                       const int32_t *filter_dev, double dtime, int32_t plain)
{
    if (!ctx || check_layout(lay) || !col || n < 0 || (!filter_dev && n > lay->ldc)) return CTSM_E_ARG;
    if (n == 0) return CTSM_OK;
    if (plain) {
        int rc;
        if ((rc = ensure_work(ctx, soilwater_plain_workspace_doubles(n)))) return rc;
        launch_soilwater_plain(ctx->stream, make_view(ctx, *lay), ctx->prm, *col, n, filter_dev, dtime,
                               (double *)ctx->work.p, n, ctx->err);
    } else {
        const View V = make_view(ctx, *lay);
        const int32_t *rf = filter_dev;
        {                                        // columns sorted by their previous sub-step count
            int rc;
            if ((rc = ensure_work(ctx, (richards_order_bytes(n) + 7) / 8))) return rc;
            rf = launch_richards_order(ctx->stream, V, *col, n, filter_dev, ctx->work.p, ctx->tun.richards_lanes);
        }
        launch_soilwater(ctx->stream, V, ctx->prm, *col, n, rf, dtime, ctx->err, ctx->tun.richards_lanes,
                         ctx->tun.richards_block);
    }
    return finish_launch(ctx);
}

int ctsm_soiltemperature_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                             const ctsm_patch_state *pch, int32_t n, const int32_t *filter_dev, double dtime,
                             int32_t plain)
{
    if (!ctx || check_layout(lay) || !col || !pch || n < 0 || (!filter_dev && n > lay->ldc)) return CTSM_E_ARG;
    if (n == 0) return CTSM_OK;
    int rc;
    if (plain) {
        if ((rc = ensure_work(ctx, soiltemp_plain_workspace_doubles(n)))) return rc;
        launch_soiltemp_plain(ctx->stream, make_view(ctx, *lay), ctx->prm, *col, *pch, n, filter_dev, dtime,
                              (double *)ctx->work.p, n, ctx->err);
    } else {
        if ((rc = ensure_work(ctx, soiltemp_workspace_doubles(n)))) return rc;
        launch_soiltemp(ctx->stream, make_view(ctx, *lay), ctx->prm, *col, *pch, n, filter_dev, dtime,
                        (double *)ctx->work.p, n, ctx->err);
    }
    return finish_launch(ctx);
}

int ctsm_soilwater_fractions_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col, int32_t n,
                                 const int32_t *filter_dev)
{
    if (!ctx || check_layout(lay) || !col || n < 0 || (!filter_dev && n > lay->ldc)) return CTSM_E_ARG;
    if (n == 0) return CTSM_OK;
    launch_soilwater_fractions(ctx->stream, make_view(ctx, *lay), *col, n, filter_dev);
    return finish_launch(ctx);
}

// Streams a fused soil step over n columns is split across (1 = not split).
static int step_split(const ctsm_ctx *ctx, int n)
{
    const int S = ctx->tun.step_streams;
    return (!ctx->tun.fused_single_kernel && S > 1 && n >= 64 * S) ? S : 1;
}
// Workspace doubles and per-part filter bytes of a fused soil step.
static size_t step_work_doubles(const ctsm_ctx *ctx, int n)
{
    // per part the heat solve's rows, reused for the Richards order after it
    const int S = step_split(ctx, n);
    const int len = (n + S - 1) / S;
    return (size_t)S * std::max(soiltemp_workspace_doubles(len), (richards_order_bytes(len) + 7) / 8);
}
// Workspace doubles of a full land step (one stream): the heat solve's LU
// rows, or in the plain form the larger of the plain kernels' intermediates.
static size_t land_work_doubles(const ctsm_ctx *ctx, int n)
{
    if (ctx->tun.plain_routines)
        return std::max(soiltemp_plain_workspace_doubles(n), soilwater_plain_workspace_doubles(n));
    // the heat solve's rows, then (reused once the heat solve is done) the
    // Richards solve's column order
    return std::max(soiltemp_workspace_doubles(n), (richards_order_bytes(n) + 7) / 8);
}
static size_t step_partf_bytes(const ctsm_ctx *ctx, int n)
{
    return step_split(ctx, n) == 1 ? 0 : (size_t)n * sizeof(int32_t);
}

// nsteps fused soil steps.  Split over S streams, each part runs all its
// steps on its own stream between one fork and one join, so the parts drift
// apart and the heat solve of one part (one wave per SIMD, latency-bound)
// overlaps the Richards solve of another; results are per column, so
// identical to the unsplit steps.  Parts are contiguous slices of the filter
// so a wave's columns stay adjacent in memory; a caller that stores its batch
// binned by work keeps the parts balanced by dealing the bins over them.
static int soil_steps(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col, const ctsm_patch_state *pch,
                      int32_t n, const int32_t *filter_dev, double dtime, int nsteps)
{
    if (!ctx || check_layout(lay) || !col || !pch || n < 0 || nsteps < 1 || (!filter_dev && n > lay->ldc))
        return CTSM_E_ARG;
    if (n == 0) return CTSM_OK;
    int rc;
    if ((rc = ensure_work(ctx, step_work_doubles(ctx, n)))) return rc;
    const View V = make_view(ctx, *lay);
    const int S = step_split(ctx, n);
    if (ctx->tun.fused_single_kernel) {
        for (int s = 0; s < nsteps; ++s)
            launch_soil_step(ctx->stream, V, ctx->prm, *col, *pch, n, filter_dev, dtime, (double *)ctx->work.p, n,
                             ctx->err);
    } else if (S > 1) {
        const long len = (n + S - 1) / S;
        hipError_t he = hipSuccess;
        const int32_t *pf = filter_dev;
        if (!pf) {
            he = ctx->partf.ensure(step_partf_bytes(ctx, n));
            if (he != hipSuccess) return hip_fail(ctx, he);
            hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, lay->begc_alloc, n,
                               (int32_t *)ctx->partf.p);
            pf = (const int32_t *)ctx->partf.p;
        }
        he = hipEventRecord(ctx->ev_fork, ctx->stream);
        for (int k = 1; he == hipSuccess && k < S; ++k) he = hipStreamWaitEvent(ctx->side[k - 1], ctx->ev_fork, 0);
        if (he != hipSuccess) return hip_fail(ctx, he);
        for (int k = 0; k < S; ++k) {
            hipStream_t st = k == 0 ? ctx->stream : ctx->side[k - 1];
            const int cnt = (int)std::min<long>(len, (long)n - k * len);
            if (cnt <= 0) continue;
            double *ws = (double *)ctx->work.p + (size_t)k * (step_work_doubles(ctx, n) / S);
            for (int s = 0; s < nsteps; ++s) {
                launch_soiltemp_fractions(st, V, ctx->prm, *col, *pch, cnt, pf + k * len, dtime, ws, len, ctx->err,
                                          ctx->tun.soiltemp_block);
                const int32_t *rf = pf + k * len;
                rf = launch_richards_order(st, V, *col, cnt, rf, ws, ctx->tun.richards_lanes);
                launch_soilwater(st, V, ctx->prm, *col, cnt, rf, dtime, ctx->err, ctx->tun.richards_lanes,
                                 ctx->tun.richards_block);
            }
        }
        for (int k = 1; he == hipSuccess && k < S; ++k) {
            he = hipEventRecord(ctx->ev_join[k - 1], ctx->side[k - 1]);
            if (he == hipSuccess) he = hipStreamWaitEvent(ctx->stream, ctx->ev_join[k - 1], 0);
        }
        if (he != hipSuccess) return hip_fail(ctx, he);
    } else {
        for (int s = 0; s < nsteps; ++s) {
            launch_soiltemp_fractions(ctx->stream, V, ctx->prm, *col, *pch, n, filter_dev, dtime,
                                      (double *)ctx->work.p, n, ctx->err, ctx->tun.soiltemp_block);
            const int32_t *rf = filter_dev;
            rf = launch_richards_order(ctx->stream, V, *col, n, rf, ctx->work.p, ctx->tun.richards_lanes);
            launch_soilwater(ctx->stream, V, ctx->prm, *col, n, rf, dtime, ctx->err, ctx->tun.richards_lanes,
                             ctx->tun.richards_block);
        }
    }
    return finish_launch(ctx);
}

int ctsm_soil_step_fused_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                             const ctsm_patch_state *pch, int32_t n, const int32_t *filter_dev, double dtime)
{
    return soil_steps(ctx, lay, col, pch, n, filter_dev, dtime, 1);
}

int ctsm_soil_steps_fused_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                              const ctsm_patch_state *pch, int32_t n, const int32_t *filter_dev, double dtime,
                              int32_t nsteps)
{
    return soil_steps(ctx, lay, col, pch, n, filter_dev, dtime, nsteps);
}

// ---- canopy ----------------------------------------------------------------

int ctsm_set_canopy_params(ctsm_ctx *ctx, const ctsm_canopy_params *prm, const ctsm_pft_params *pft)
{
    if (!ctx || !prm || !pft || pft->npft < 1 || pft->npft > CTSM_MAX_PFT) return CTSM_E_ARG;
    const double *cols[] = {pft->dleaf, pft->c3psn, pft->slatop, pft->leafcn, pft->flnr, pft->fnitr, pft->mbbopt,
                            pft->medlynslope, pft->medlynintercept, pft->smpso, pft->smpsc, pft->theta_cj,
                            pft->z0v_Cr, pft->z0v_Cs, pft->z0v_c, pft->z0v_cw, pft->z0v_LAImax};
    const int ncols = (int)(sizeof(cols) / sizeof(cols[0]));
    const double *cols2[] = {pft->dbh, pft->fbw, pft->nstem, pft->rstem_per_dbh, pft->wood_density};
    for (const double *a : cols) if (!a) return CTSM_E_ARG;
    for (const double *a : cols2) if (!a) return CTSM_E_ARG;
    if (!pft->is_tree || !pft->is_shrub) return CTSM_E_ARG;
    // table rows in the order of the kernel's field enum
    std::vector<double> tab((size_t)24 * CTSM_MAX_PFT, 0.0);
    auto put = [&](int field, const double *a) {
        for (int i = 0; i < pft->npft; ++i) tab[(size_t)field * CTSM_MAX_PFT + i] = a[i];
    };
    for (int f = 0; f < ncols; ++f) put(f, cols[f]);
    for (int i = 0; i < pft->npft; ++i) {
        tab[(size_t)17 * CTSM_MAX_PFT + i] = (double)pft->is_tree[i];
        tab[(size_t)18 * CTSM_MAX_PFT + i] = (double)pft->is_shrub[i];
    }
    for (int f = 0; f < 5; ++f) put(19 + f, cols2[f]);
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e);
    if ((e = ctx->pft.ensure(tab.size() * sizeof(double))) != hipSuccess) return hip_fail(ctx, e);
    if ((e = hipMemcpyAsync(ctx->pft.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                            ctx->stream)) != hipSuccess)
        return hip_fail(ctx, e);
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e);
    ctx->cprm = *prm;
    ctx->have_canopy = true;
    return CTSM_OK;
}

static int check_canopy_layout(const ctsm_ctx *ctx, const ctsm_layout *L)
{
    if (!ctx->have_canopy) return CTSM_E_ARG;
    if (check_layout(L) || L->ldp <= 0 || L->ldg <= 0) return CTSM_E_ARG;
    return CTSM_OK;
}

// Grows the canopy round buffer to n patches and launches CanopyFluxes.
// With fixed_work, the rounds use that buffer of fixed_bytes instead and
// nothing is allocated (CTSM_E_SCRATCH when it is too small).  With
// bounds_are_filter the per-step resets of the bounds (rb1, the
// photosynthesis outputs) cover the filter's patches instead of the whole
// layout: a caller whose layout holds other, independent runs (the ensemble
// server's pool) must not touch their patches.
static int run_canopy(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state &col, const ctsm_canopy_state &can,
                      int32_t n, const int32_t *filter_dev, double dtime, bool zero_cgrnd,
                      void *fixed_work = nullptr, size_t fixed_bytes = 0, bool bounds_are_filter = false,
                      const ctsm_phs_state *phs_q = nullptr, const ctsm_phs_canopy_state *phs_h = nullptr)
{
    // the configuration fixes the form: the PHS one only through the ctsm_canopyfluxes_phs entries
    if ((ctx->cfg.use_hydrstress != 0) != (phs_h != nullptr)) return CTSM_E_CONFIG;
    const bool one_pass = ctx->tun.plain_routines != 0;
    const int k0 = one_pass ? 1001 : ctx->tun.canopy_first_iterations;
    const int k = one_pass ? 1 : ctx->tun.canopy_round_iterations;
    const size_t wb = canopy_work_bytes(n, ctx->cfg.itmax_canopy_fluxes, k0, k);
    void *work = nullptr;
    if (wb && fixed_work) {
        if (wb > fixed_bytes) {
            ctx->last_code = CTSM_E_SCRATCH;
            ctx->last_index = (int32_t)std::min<size_t>((wb + 1023) / 1024, INT32_MAX);
            return CTSM_E_SCRATCH;
        }
        work = fixed_work;
    } else if (wb) {
        hipError_t he = ctx->cwork.ensure(wb);
        if (he != hipSuccess) return hip_fail(ctx, he);
        work = ctx->cwork.p;
    }
    launch_canopy(ctx->stream, *lay, ctx->cfg, ctx->cprm, (const double *)ctx->pft.p, col, can, n, filter_dev, dtime,
                  ctx->err, work, k0, k, zero_cgrnd, ctx->tun.canopy_waves, bounds_are_filter, phs_q, phs_h,
                  (const double *)ctx->phs_tab.p);
    return CTSM_OK;
}

int ctsm_canopyfluxes_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                          const ctsm_canopy_state *can, int32_t n, const int32_t *filterp_dev, double dtime)
{
    if (!ctx || !col || !can || n < 0 || check_canopy_layout(ctx, lay) || (!filterp_dev && n > lay->ldp))
        return CTSM_E_ARG;
    int rc = run_canopy(ctx, lay, *col, *can, n, filterp_dev, dtime, false);
    if (rc) return rc;
    return finish_launch(ctx);
}

int ctsm_canopyfluxes_dev_noalloc(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                                  const ctsm_canopy_state *can, int32_t n, const int32_t *filterp_dev, double dtime)
{
    if (!ctx || !col || !can || n < 0 || check_canopy_layout(ctx, lay) || (!filterp_dev && n > lay->ldp))
        return CTSM_E_ARG;
    if (!ctx->scratch_dev) return CTSM_E_HIP;
    int rc = run_canopy(ctx, lay, *col, *can, n, filterp_dev, dtime, false, ctx->scratch_dev,
                        CTSM_NOALLOC_SCRATCH_BYTES);
    if (rc) return rc;
    return finish_launch(ctx);
}

int ctsm_root_sink_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                       const ctsm_canopy_state *can, int32_t n, const int32_t *filterc_dev)
{
    if (!ctx || !col || !can || n < 0 || check_layout(lay) || (!filterc_dev && n > lay->ldc)) return CTSM_E_ARG;
    launch_root_sink(ctx->stream, *lay, ctx->cfg, *col, *can, n, filterc_dev, ctx->err);
    return finish_launch(ctx);
}

int ctsm_land_step_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                       const ctsm_patch_state *pch, const ctsm_canopy_state *can, int32_t nc,
                       const int32_t *filterc_dev, int32_t np, const int32_t *filterp_dev, double dtime)
{
    if (!ctx || !col || !pch || !can || nc < 0 || np < 0 || check_canopy_layout(ctx, lay) ||
        (!filterc_dev && nc > lay->ldc) || (!filterp_dev && np > lay->ldp))
        return CTSM_E_ARG;
    // the arrays CanopyFluxes writes and the soil solves read must coincide
    if (pch->dlrad != can->dlrad || pch->cgrnd != can->cgrnd || pch->eflx_sh_grnd != can->eflx_sh_grnd ||
        pch->eflx_sh_snow != can->eflx_sh_snow || pch->eflx_sh_soil != can->eflx_sh_soil ||
        pch->eflx_sh_h2osfc != can->eflx_sh_h2osfc || pch->qflx_evap_soi != can->qflx_evap_soi ||
        pch->qflx_ev_snow != can->qflx_ev_snow || pch->qflx_ev_soil != can->qflx_ev_soil ||
        pch->qflx_ev_h2osfc != can->qflx_ev_h2osfc || col->qflx_rootsoi != can->qflx_rootsoi_col)
        return CTSM_E_ARG;
    int rc = run_canopy(ctx, lay, *col, *can, np, filterp_dev, dtime, true);
    if (rc) return rc;
    launch_root_sink(ctx->stream, *lay, ctx->cfg, *col, *can, nc, filterc_dev, ctx->err);
    rc = finish_launch(ctx);
    if (rc) return rc;
    return ctsm_soil_step_fused_dev(ctx, lay, col, pch, nc, filterc_dev, dtime);
}

// ---- captured step sequences ---------------------------------------------

struct ctsm_graph {
    ctsm_ctx *ctx = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t ex = nullptr;
    DevBuf work, cwork, partf;   // workspaces the captured kernels use
};

int ctsm_graph_destroy(ctsm_graph *gr)
{
    if (!gr) return CTSM_OK;
    if (gr->ctx) {
        (void)hipSetDevice(gr->ctx->device);
        (void)hipStreamSynchronize(gr->ctx->stream);
    }
    if (gr->ex) (void)hipGraphExecDestroy(gr->ex);
    if (gr->g) (void)hipGraphDestroy(gr->g);
    gr->work.release();
    gr->cwork.release();
    gr->partf.release();
    delete gr;
    return CTSM_OK;
}

// Captures what `steps` enqueues into a graph with its own workspaces (sized
// before the capture: nothing may allocate under it).
extern "C++" {
template <class Steps>
static int capture_graph(ctsm_ctx *ctx, int32_t nc, int32_t np, bool canopy, ctsm_graph **out, Steps steps)
{
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    ctsm_graph *gr = new ctsm_graph;
    gr->ctx = ctx;
    he = gr->work.ensure(std::max<size_t>(std::max(step_work_doubles(ctx, nc), land_work_doubles(ctx, nc)), 1) *
                         sizeof(double));
    if (he == hipSuccess && step_partf_bytes(ctx, nc)) he = gr->partf.ensure(step_partf_bytes(ctx, nc));
    const size_t cwb = canopy ? canopy_work_bytes(np, ctx->cfg.itmax_canopy_fluxes, ctx->tun.canopy_first_iterations,
                                                  ctx->tun.canopy_round_iterations)
                              : 0;
    if (he == hipSuccess && cwb) he = gr->cwork.ensure(cwb);
    if (he != hipSuccess) {
        ctsm_graph_destroy(gr);
        return hip_fail(ctx, he);
    }
    std::swap(ctx->work, gr->work);
    std::swap(ctx->cwork, gr->cwork);
    std::swap(ctx->partf, gr->partf);
    int rc = CTSM_OK;
    he = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (he == hipSuccess) {
        rc = steps();
        hipError_t he2 = hipStreamEndCapture(ctx->stream, &gr->g);
        if (he2 != hipSuccess && rc == CTSM_OK) rc = hip_fail(ctx, he2);
    } else {
        rc = hip_fail(ctx, he);
    }
    std::swap(ctx->work, gr->work);
    std::swap(ctx->cwork, gr->cwork);
    std::swap(ctx->partf, gr->partf);
    if (rc == CTSM_OK) {
        he = hipGraphInstantiate(&gr->ex, gr->g, nullptr, nullptr, 0);
        if (he != hipSuccess) rc = hip_fail(ctx, he);
    }
    if (rc != CTSM_OK) {
        gr->ctx = nullptr;      // nothing of it was launched
        ctsm_graph_destroy(gr);
        return rc;
    }
    *out = gr;
    return CTSM_OK;
}
}  // extern "C++"

int ctsm_graph_capture(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                       const ctsm_patch_state *pch, const ctsm_canopy_state *can, int32_t nc,
                       const int32_t *filterc_dev, int32_t np, const int32_t *filterp_dev, double dtime,
                       int32_t nsteps, ctsm_graph **out)
{
    if (!out) return CTSM_E_ARG;
    *out = nullptr;
    if (!ctx || !col || !pch || nc < 0 || np < 0 || nsteps < 1 || check_layout(lay) ||
        (!filterc_dev && nc > lay->ldc) || (can && (check_canopy_layout(ctx, lay) || (!filterp_dev && np > lay->ldp))))
        return CTSM_E_ARG;
    return capture_graph(ctx, nc, np, can != nullptr, out, [&]() {
        if (!can) return soil_steps(ctx, lay, col, pch, nc, filterc_dev, dtime, nsteps);
        int rc = CTSM_OK;
        for (int s = 0; s < nsteps && rc == CTSM_OK; ++s)
            rc = ctsm_land_step_dev(ctx, lay, col, pch, can, nc, filterc_dev, np, filterp_dev, dtime);
        return rc;
    });
}

int ctsm_graph_capture_full(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                            const ctsm_patch_state *pch, const ctsm_canopy_state *can,
                            const ctsm_soilflux_state *sf, const ctsm_surface_water_state *sw,
                            const ctsm_hydro_params *hprm, const ctsm_water_table_state *wt,
                            const ctsm_drainage_state *dr, int32_t nc, const int32_t *filterc_dev, int32_t np,
                            const int32_t *filterp_dev, double dtime, int32_t nsteps, ctsm_graph **out)
{
    if (!out) return CTSM_E_ARG;
    *out = nullptr;
    if (!ctx || !col || !pch || !can || !sf || nc < 0 || np < 0 || nsteps < 1 || check_canopy_layout(ctx, lay) ||
        (!filterc_dev && nc > lay->ldc) || (!filterp_dev && np > lay->ldp))
        return CTSM_E_ARG;
    return capture_graph(ctx, nc, np, true, out, [&]() {
        int rc = CTSM_OK;
        for (int s = 0; s < nsteps && rc == CTSM_OK; ++s)
            rc = ctsm_land_step_full_dev(ctx, lay, col, pch, can, sf, sw, hprm, wt, dr, nc, filterc_dev, np,
                                         filterp_dev, dtime);
        return rc;
    });
}

int ctsm_graph_launch(ctsm_ctx *ctx, ctsm_graph *gr)
{
    if (!ctx || !gr || gr->ctx != ctx || !gr->ex) return CTSM_E_ARG;
    hipError_t he = hipGraphLaunch(gr->ex, ctx->stream);
    return he == hipSuccess ? CTSM_OK : hip_fail(ctx, he);
}

static int canopy_mirror(ctsm_ctx *ctx, const ctsm_layout *L, const ctsm_canopy_state *can, ctsm_canopy_state *dcan,
                         bool to_device)
{
    int rc = mirror(ctx, kCanopyPatch, kNCanP, can, dcan, ctx->canp, L->begp_alloc, L->ldp,
                    Range{L->begp_alloc, L->begp_alloc + L->ldp - 1}, to_device);
    if (rc) return rc;
    rc = mirror(ctx, kCanopyCol, kNCanC, can, dcan, ctx->canc, L->begc_alloc, L->ldc,
                Range{L->begc_alloc, L->begc_alloc + L->ldc - 1}, to_device);
    if (rc) return rc;
    return mirror(ctx, kCanopyGrid, kNCanG, can, dcan, ctx->cang, L->begg_alloc, L->ldg,
                  Range{L->begg_alloc, L->begg_alloc + L->ldg - 1}, to_device);
}

static int upload_filter(ctsm_ctx *ctx, int32_t n, const int32_t *filter, const int32_t **dfilt)
{
    *dfilt = nullptr;
    if (!filter || n == 0) return CTSM_OK;
    hipError_t he = ctx->filt.ensure((size_t)n * sizeof(int32_t));
    if (he != hipSuccess) return hip_fail(ctx, he);
    he = hipMemcpyAsync(ctx->filt.p, filter, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he);
    *dfilt = (const int32_t *)ctx->filt.p;
    return CTSM_OK;
}

// ---- SoilFluxes and the soil energy balance check ---------------------------

static bool sf_complete(const ctsm_soilflux_state *sf)
{
    for (int f = 0; f < kNSfP; ++f)
        if (!*(void *const *)((const char *)sf + kSfPatch[f].offset)) return false;
    for (int f = 0; f < kNSfC; ++f)
        if (!*(void *const *)((const char *)sf + kSfCol[f].offset)) return false;
    return true;
}

int ctsm_soilfluxes_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                        const ctsm_patch_state *pch, const ctsm_soilflux_state *sf, int32_t nc,
                        const int32_t *filterc_dev, int32_t np, const int32_t *filterp_dev, double dtime)
{
    if (!ctx || !col || !pch || !sf || nc < 0 || np < 0 || check_layout(lay) || lay->ldp <= 0 ||
        (!filterc_dev && nc > lay->ldc) || (!filterp_dev && np > lay->ldp) || !sf_complete(sf))
        return CTSM_E_ARG;
    launch_soilfluxes(ctx->stream, *lay, ctx->cfg, *col, *pch, *sf, nc, filterc_dev, np, filterp_dev, dtime, ctx->err);
    return finish_launch(ctx);
}

int ctsm_soil_balance_check_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_soilflux_state *sf, int32_t n,
                                const int32_t *filter_dev, double *errsoi_max, int32_t *index)
{
    if (!ctx || !sf || !sf->errsoi_col || !errsoi_max || !index || n < 0 || check_layout(lay) ||
        (!filter_dev && n > lay->ldc))
        return CTSM_E_ARG;
    hipError_t he = ctx->bal.ensure(2 * sizeof(unsigned long long));
    if (he != hipSuccess) return hip_fail(ctx, he);
    launch_balance(ctx->stream, *lay, ctx->cfg, sf->errsoi_col, n, filter_dev, (unsigned long long *)ctx->bal.p);
    int rc = finish_launch(ctx);
    if (rc) return rc;
    unsigned long long w[2] = {0ull, 0ull};
    he = hipMemcpyAsync(w, ctx->bal.p, sizeof(w), hipMemcpyDeviceToHost, ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he);
    double m;
    std::memcpy(&m, &w[0], sizeof(m));
    *errsoi_max = m;
    *index = (n > 0 && w[1] != 0x7fffffffull) ? (int32_t)w[1] : 0;
    return CTSM_OK;
}

int ctsm_energy_balance_check_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                                  const ctsm_patch_state *pch, const ctsm_canopy_state *can,
                                  const ctsm_soilflux_state *sf, int32_t n, const int32_t *filterp_dev, double *errseb,
                                  double *errlon, double *errseb_max, int32_t *seb_patch, double *errlon_max,
                                  int32_t *lon_patch)
{
    if (!ctx || !col || !pch || !can || !sf || !errseb_max || !seb_patch || !errlon_max || !lon_patch || n < 0 ||
        check_layout(lay) || lay->ldp <= 0 || (!filterp_dev && n > lay->ldp) || !col->forc_lwrad || !pch->sabg_chk ||
        !can->sabv || !can->dhsdt_canopy || !can->active || !can->column || !sf->eflx_lwrad_out ||
        !sf->eflx_lwrad_net || !sf->eflx_sh_tot || !sf->eflx_lh_tot || !sf->eflx_soil_grnd)
        return CTSM_E_ARG;
    hipError_t he = ctx->bal.ensure(4 * sizeof(unsigned long long));
    if (he != hipSuccess) return hip_fail(ctx, he);
    launch_energy_balance(ctx->stream, *lay, ctx->cfg, *col, *pch, *can, *sf, n, filterp_dev, errseb, errlon,
                          (unsigned long long *)ctx->bal.p, ctx->err);
    int rc = finish_launch(ctx);
    if (rc) return rc;
    unsigned long long w[4] = {0ull, 0ull, 0ull, 0ull};
    he = hipMemcpyAsync(w, ctx->bal.p, sizeof(w), hipMemcpyDeviceToHost, ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he);
    if ((rc = sync_and_check(ctx))) return rc;
    std::memcpy(errseb_max, &w[0], sizeof(double));
    std::memcpy(errlon_max, &w[2], sizeof(double));
    *seb_patch = (n > 0 && w[1] != 0x7fffffffull) ? (int32_t)w[1] : 0;
    *lon_patch = (n > 0 && w[3] != 0x7fffffffull) ? (int32_t)w[3] : 0;
    return CTSM_OK;
}

int ctsm_solar_balance_check_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_canopy_state *can, int32_t n,
                                 const int32_t *filterp_dev, const double *fsa, const double *fsr,
                                 const double *forc_solad_col, const double *forc_solai, double *errsol,
                                 double *errsol_max, int32_t *sol_patch)
{
    if (!ctx || !can || !fsa || !fsr || !forc_solad_col || !forc_solai || !errsol_max || !sol_patch || n < 0 ||
        check_layout(lay) || lay->ldp <= 0 || lay->ldg <= 0 || (!filterp_dev && n > lay->ldp) || !can->active ||
        !can->column || !can->gridcell)
        return CTSM_E_ARG;
    hipError_t he = ctx->bal.ensure(2 * sizeof(unsigned long long));
    if (he != hipSuccess) return hip_fail(ctx, he);
    launch_solar_balance(ctx->stream, *lay, ctx->cfg, *can, n, filterp_dev, fsa, fsr, forc_solad_col, forc_solai,
                         errsol, (unsigned long long *)ctx->bal.p, ctx->err);
    int rc = finish_launch(ctx);
    if (rc) return rc;
    unsigned long long w[2] = {0ull, 0ull};
    he = hipMemcpyAsync(w, ctx->bal.p, sizeof(w), hipMemcpyDeviceToHost, ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he);
    if ((rc = sync_and_check(ctx))) return rc;
    std::memcpy(errsol_max, &w[0], sizeof(double));
    *sol_patch = (n > 0 && w[1] != 0x7fffffffull) ? (int32_t)w[1] : 0;
    return CTSM_OK;
}

int ctsm_water_mass_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                        const ctsm_canopy_state *can, const ctsm_water_balance_state *wb, int32_t n,
                        const int32_t *filter_dev, double aquifer_water_baseline, int32_t subtract_dynbal_baselines,
                        double *water_mass)
{
    if (!ctx || !col || !can || !wb || !water_mass || n < 0 || check_layout(lay) || lay->ldp <= 0 ||
        (!filter_dev && n > lay->ldc) || !col->snl || !col->patchi || !col->patchf || !col->h2osoi_liq ||
        !col->h2osoi_ice || !col->h2osfc || !col->h2osno_no_layers || !can->active || !can->wtcol || !can->liqcan ||
        !can->snocan || !wb->total_plant_stored_h2o || !wb->wa || !wb->hydrologically_active ||
        (subtract_dynbal_baselines && (!wb->dynbal_baseline_liq || !wb->dynbal_baseline_ice)))
        return CTSM_E_ARG;
    launch_water_mass(ctx->stream, *lay, ctx->cfg, *col, *can, *wb, n, filter_dev, aquifer_water_baseline,
                      subtract_dynbal_baselines ? 1 : 0, water_mass, ctx->err);
    return finish_launch(ctx);
}

int ctsm_water_balance_check_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_water_balance_state *wb,
                                 int32_t n, const int32_t *filterc_dev, double dtime, double *errh2o_max,
                                 int32_t *errh2o_col)
{
    if (!ctx || !wb || !errh2o_max || !errh2o_col || n < 0 || check_layout(lay) ||
        (!filterc_dev && n > lay->ldc) || !wb->col_active || !wb->begwb || !wb->endwb || !wb->forc_rain ||
        !wb->forc_snow || !wb->qflx_floodc || !wb->qflx_sfc_irrig || !wb->qflx_glcice_dyn_water_flux ||
        !wb->qflx_evap_tot || !wb->qflx_surf || !wb->qflx_qrgwl || !wb->qflx_drain || !wb->qflx_drain_perched ||
        !wb->qflx_ice_runoff || !wb->qflx_snwcp_discarded_liq || !wb->qflx_snwcp_discarded_ice)
        return CTSM_E_ARG;
    hipError_t he = ctx->bal.ensure(2 * sizeof(unsigned long long));
    if (he != hipSuccess) return hip_fail(ctx, he);
    launch_water_balance(ctx->stream, *lay, ctx->cfg, *wb, n, filterc_dev, dtime, (unsigned long long *)ctx->bal.p,
                         ctx->err);
    int rc = finish_launch(ctx);
    if (rc) return rc;
    unsigned long long w[2] = {0ull, 0ull};
    he = hipMemcpyAsync(w, ctx->bal.p, sizeof(w), hipMemcpyDeviceToHost, ctx->stream);
    if (he != hipSuccess) return hip_fail(ctx, he);
    if ((rc = sync_and_check(ctx))) return rc;
    std::memcpy(errh2o_max, &w[0], sizeof(double));
    *errh2o_col = (n > 0 && w[1] != 0x7fffffffull) ? (int32_t)w[1] : 0;
    return CTSM_OK;
}

int ctsm_soilfluxes(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col, const ctsm_patch_state *pch,
                    const ctsm_soilflux_state *sf, int32_t nc, const int32_t *filterc, int32_t np,
                    const int32_t *filterp, double dtime)
{
    if (!ctx || !col || !pch || !sf || nc < 0 || np < 0 || check_layout(lay) || lay->ldp <= 0 ||
        (!filterc && nc > lay->ldc) || (!filterp && np > lay->ldp) || !sf_complete(sf))
        return CTSM_E_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    for (int i = 0; filterc && i < nc; ++i)
        if (filterc[i] < lay->begc_alloc || filterc[i] >= lay->begc_alloc + lay->ldc) return CTSM_E_ARG;
    for (int i = 0; filterp && i < np; ++i)
        if (filterp[i] < lay->begp_alloc || filterp[i] >= lay->begp_alloc + lay->ldp) return CTSM_E_ARG;
    // only the members SoilFluxes reads are mirrored
    ctsm_col_state hc{};
    hc.snl = col->snl; hc.patchi = col->patchi; hc.patchf = col->patchf;
    hc.frac_sno_eff = col->frac_sno_eff; hc.frac_h2osfc = col->frac_h2osfc;
    hc.h2osoi_liq = col->h2osoi_liq; hc.h2osoi_ice = col->h2osoi_ice; hc.t_soisno = col->t_soisno;
    hc.t_h2osfc = col->t_h2osfc; hc.t_grnd = col->t_grnd; hc.emg = col->emg; hc.forc_lwrad = col->forc_lwrad;
    hc.htvp = col->htvp; hc.xmf = col->xmf; hc.xmf_h2osfc = col->xmf_h2osfc; hc.fact = col->fact;
    hc.c_h2osfc = col->c_h2osfc; hc.eflx_h2osfc_to_snow = col->eflx_h2osfc_to_snow;
    if (!hc.snl || !hc.patchi || !hc.patchf || !hc.frac_sno_eff || !hc.frac_h2osfc || !hc.h2osoi_liq ||
        !hc.h2osoi_ice || !hc.t_soisno || !hc.t_h2osfc || !hc.t_grnd || !hc.emg || !hc.forc_lwrad || !hc.htvp ||
        !hc.xmf || !hc.xmf_h2osfc || !hc.fact || !hc.c_h2osfc || !hc.eflx_h2osfc_to_snow)
        return CTSM_E_ARG;
    ctsm_patch_state hp{};
    hp.column = pch->column; hp.wtcol = pch->wtcol; hp.frac_veg_nosno = pch->frac_veg_nosno;
    hp.sabg_soil = pch->sabg_soil; hp.sabg_snow = pch->sabg_snow; hp.dlrad = pch->dlrad;
    if (!hp.column || !hp.wtcol || !hp.frac_veg_nosno || !hp.sabg_soil || !hp.sabg_snow || !hp.dlrad)
        return CTSM_E_ARG;
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    const Range pr{lay->begp_alloc, lay->begp_alloc + lay->ldp - 1};
    ctsm_col_state dcol{};
    ctsm_patch_state dpch{};
    ctsm_soilflux_state dsf{};
    const int ncf = (int)(sizeof(kColFields) / sizeof(kColFields[0]));
    const int npf = (int)(sizeof(kPatchFields) / sizeof(kPatchFields[0]));
    int rc = mirror(ctx, kColFields, ncf, &hc, &dcol, ctx->colmir, lay->begc_alloc, lay->ldc, cr, true);
    if (!rc) rc = mirror(ctx, kPatchFields, npf, &hp, &dpch, ctx->pchmir, lay->begp_alloc, lay->ldp, pr, true);
    if (!rc) rc = mirror(ctx, kSfPatch, kNSfP, sf, &dsf, ctx->sfp, lay->begp_alloc, lay->ldp, pr, true);
    if (!rc) rc = mirror(ctx, kSfCol, kNSfC, sf, &dsf, ctx->sfc, lay->begc_alloc, lay->ldc, cr, true);
    if (rc) return rc;
    const int32_t *dfc = nullptr, *dfp = nullptr;
    if ((rc = upload_filter(ctx, nc, filterc, &dfc))) return rc;
    if (filterp && np > 0) {
        he = ctx->filt2.ensure((size_t)np * sizeof(int32_t));
        if (he == hipSuccess)
            he = hipMemcpyAsync(ctx->filt2.p, filterp, (size_t)np * sizeof(int32_t), hipMemcpyHostToDevice,
                                ctx->stream);
        if (he != hipSuccess) return hip_fail(ctx, he);
        dfp = (const int32_t *)ctx->filt2.p;
    }
    launch_soilfluxes(ctx->stream, *lay, ctx->cfg, dcol, dpch, dsf, nc, dfc, np, dfp, dtime, ctx->err);
    if ((rc = finish_launch(ctx))) return rc;
    rc = mirror(ctx, kSfPatch, kNSfP, sf, &dsf, ctx->sfp, lay->begp_alloc, lay->ldp, pr, false);
    if (!rc) rc = mirror(ctx, kSfCol, kNSfC, sf, &dsf, ctx->sfc, lay->begc_alloc, lay->ldc, cr, false);
    if (rc) return rc;
    return sync_and_check(ctx);
}

int ctsm_soil_balance_check(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_soilflux_state *sf, int32_t n,
                            const int32_t *filterc, double *errsoi_max, int32_t *index)
{
    if (!ctx || !sf || !sf->errsoi_col || !errsoi_max || !index || n < 0 || check_layout(lay) ||
        (!filterc && n > lay->ldc))
        return CTSM_E_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    for (int i = 0; filterc && i < n; ++i)
        if (filterc[i] < lay->begc_alloc || filterc[i] >= lay->begc_alloc + lay->ldc) return CTSM_E_ARG;
    ctsm_soilflux_state hs{};
    hs.errsoi_col = sf->errsoi_col;
    ctsm_soilflux_state dsf{};
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    int rc = mirror(ctx, kSfCol, kNSfC, &hs, &dsf, ctx->sfc, lay->begc_alloc, lay->ldc, cr, true);
    if (rc) return rc;
    const int32_t *dfc = nullptr;
    if ((rc = upload_filter(ctx, n, filterc, &dfc))) return rc;
    return ctsm_soil_balance_check_dev(ctx, lay, &dsf, n, dfc, errsoi_max, index);
}

static bool sw_complete(const ctsm_surface_water_state *sw)
{
    for (int f = 0; f < kNSw; ++f)
        if (!*(void *const *)((const char *)sw + kSw[f].offset)) return false;
    return true;
}

int ctsm_surface_water_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                           const ctsm_surface_water_state *sw, const ctsm_hydro_params *prm, int32_t n,
                           const int32_t *filter_dev, double dtime)
{
    if (!ctx || !col || !sw || !prm || n < 0 || check_layout(lay) || (!filter_dev && n > lay->ldc) ||
        !sw_complete(sw) || !col->snl || !col->frac_sno_eff || !col->frac_h2osfc || !col->icefrac || !col->hksat)
        return CTSM_E_ARG;
    launch_surface_water(ctx->stream, *lay, ctx->cfg.e_ice, *prm, *col, *sw, n, filter_dev, dtime, ctx->err);
    return finish_launch(ctx);
}

int ctsm_surface_water(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                       const ctsm_surface_water_state *sw, const ctsm_hydro_params *prm, int32_t n,
                       const int32_t *filter, double dtime)
{
    if (!ctx || !col || !sw || !prm || n < 0 || check_layout(lay) || (!filter && n > lay->ldc) || !sw_complete(sw) ||
        !col->snl || !col->frac_sno_eff || !col->frac_h2osfc || !col->icefrac || !col->hksat)
        return CTSM_E_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    for (int i = 0; filter && i < n; ++i)
        if (filter[i] < lay->begc_alloc || filter[i] >= lay->begc_alloc + lay->ldc) return CTSM_E_ARG;
    ctsm_col_state hc{};
    hc.snl = col->snl; hc.frac_sno_eff = col->frac_sno_eff; hc.frac_h2osfc = col->frac_h2osfc;
    hc.icefrac = col->icefrac; hc.hksat = col->hksat;
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    ctsm_col_state dcol{};
    ctsm_surface_water_state dsw{};
    const int ncf = (int)(sizeof(kColFields) / sizeof(kColFields[0]));
    int rc = mirror(ctx, kColFields, ncf, &hc, &dcol, ctx->colmir, lay->begc_alloc, lay->ldc, cr, true);
    if (!rc) rc = mirror(ctx, kSw, kNSw, sw, &dsw, ctx->swm, lay->begc_alloc, lay->ldc, cr, true);
    if (rc) return rc;
    const int32_t *dfc = nullptr;
    if ((rc = upload_filter(ctx, n, filter, &dfc))) return rc;
    launch_surface_water(ctx->stream, *lay, ctx->cfg.e_ice, *prm, dcol, dsw, n, dfc, dtime, ctx->err);
    if ((rc = finish_launch(ctx))) return rc;
    if ((rc = mirror(ctx, kSw, kNSw, sw, &dsw, ctx->swm, lay->begc_alloc, lay->ldc, cr, false))) return rc;
    return sync_and_check(ctx);
}

// the members every water-table call needs (h2osoi_vol may be NULL)
static bool wt_complete(const ctsm_water_table_state *wt)
{
    return wt->zwt && wt->zwt_perched && wt->frost_table && wt->qflx_liqdew_to_top_layer &&
           wt->qflx_soliddew_to_top_layer && wt->qflx_solidevap_from_top_layer;
}

static bool wt_col_complete(const ctsm_col_state *col)
{
    return col->t_soisno && col->dz && col->z && col->zi && col->watsat && col->nbedrock && col->snl &&
           col->frac_h2osfc && col->h2osoi_liq && col->h2osoi_ice;
}

int ctsm_water_table_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                         const ctsm_water_table_state *wt, int32_t n, const int32_t *filter_dev, double dtime)
{
    if (!ctx || !col || !wt || n < 0 || check_layout(lay) || (!filter_dev && n > lay->ldc) || !wt_complete(wt) ||
        !wt_col_complete(col))
        return CTSM_E_ARG;
    if (ctx->cfg.lower_boundary_condition == CTSM_BC_WATERTABLE) return CTSM_E_CONFIG;
    launch_water_table(ctx->stream, *lay, ctx->cfg, *col, *wt, n, filter_dev, dtime, ctx->err);
    return finish_launch(ctx);
}

int ctsm_water_table(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                     const ctsm_water_table_state *wt, int32_t n, const int32_t *filter, double dtime)
{
    if (!ctx || !col || !wt || n < 0 || check_layout(lay) || (!filter && n > lay->ldc) || !wt_complete(wt) ||
        !wt_col_complete(col))
        return CTSM_E_ARG;
    if (ctx->cfg.lower_boundary_condition == CTSM_BC_WATERTABLE) return CTSM_E_CONFIG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    for (int i = 0; filter && i < n; ++i)
        if (filter[i] < lay->begc_alloc || filter[i] >= lay->begc_alloc + lay->ldc) return CTSM_E_ARG;
    ctsm_col_state hc{};
    hc.t_soisno = col->t_soisno; hc.dz = col->dz; hc.z = col->z; hc.zi = col->zi; hc.watsat = col->watsat;
    hc.nbedrock = col->nbedrock; hc.snl = col->snl; hc.frac_h2osfc = col->frac_h2osfc;
    hc.h2osoi_liq = col->h2osoi_liq; hc.h2osoi_ice = col->h2osoi_ice;
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    ctsm_col_state dcol{};
    ctsm_water_table_state dwt{};
    const int ncf = (int)(sizeof(kColFields) / sizeof(kColFields[0]));
    int rc = mirror(ctx, kColFields, ncf, &hc, &dcol, ctx->colmir, lay->begc_alloc, lay->ldc, cr, true);
    if (!rc) rc = mirror(ctx, kWt, kNWt, wt, &dwt, ctx->wtm, lay->begc_alloc, lay->ldc, cr, true);
    if (rc) return rc;
    const int32_t *dfc = nullptr;
    if ((rc = upload_filter(ctx, n, filter, &dfc))) return rc;
    launch_water_table(ctx->stream, *lay, ctx->cfg, dcol, dwt, n, dfc, dtime, ctx->err);
    if ((rc = finish_launch(ctx))) return rc;
    if ((rc = mirror(ctx, kColFields, ncf, &hc, &dcol, ctx->colmir, lay->begc_alloc, lay->ldc, cr, false))) return rc;
    if ((rc = mirror(ctx, kWt, kNWt, wt, &dwt, ctx->wtm, lay->begc_alloc, lay->ldc, cr, false))) return rc;
    return sync_and_check(ctx);
}

// the members every drainage call needs (h2osoi_vol may be NULL)
static bool dr_complete(const ctsm_drainage_state *dr)
{
    for (int f = 0; f < kNDr; ++f) {
        if (kDr[f].offset == offsetof(ctsm_drainage_state, h2osoi_vol)) continue;
        if (!*(void *const *)((const char *)dr + kDr[f].offset)) return false;
    }
    return true;
}

static bool dr_col_complete(const ctsm_col_state *col)
{
    return col->dz && col->z && col->zi && col->watsat && col->bsw && col->sucsat && col->hksat &&
           col->eff_porosity && col->nbedrock && col->h2osoi_liq && col->h2osoi_ice && col->h2osfc && col->icefrac;
}

int ctsm_drainage_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                      const ctsm_drainage_state *dr, const ctsm_hydro_params *prm, int32_t n,
                      const int32_t *filter_dev, double dtime)
{
    if (!ctx || !col || !dr || !prm || n < 0 || check_layout(lay) || (!filter_dev && n > lay->ldc) ||
        !dr_complete(dr) || !dr_col_complete(col))
        return CTSM_E_ARG;
    if (ctx->cfg.lower_boundary_condition == CTSM_BC_WATERTABLE) return CTSM_E_CONFIG;
    launch_drainage(ctx->stream, *lay, ctx->cfg, *prm, *col, *dr, n, filter_dev, dtime, ctx->err);
    return finish_launch(ctx);
}

int ctsm_drainage(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col, const ctsm_drainage_state *dr,
                  const ctsm_hydro_params *prm, int32_t n, const int32_t *filter, double dtime)
{
    if (!ctx || !col || !dr || !prm || n < 0 || check_layout(lay) || (!filter && n > lay->ldc) ||
        !dr_complete(dr) || !dr_col_complete(col))
        return CTSM_E_ARG;
    if (ctx->cfg.lower_boundary_condition == CTSM_BC_WATERTABLE) return CTSM_E_CONFIG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    for (int i = 0; filter && i < n; ++i)
        if (filter[i] < lay->begc_alloc || filter[i] >= lay->begc_alloc + lay->ldc) return CTSM_E_ARG;
    ctsm_col_state hc{};
    hc.dz = col->dz; hc.z = col->z; hc.zi = col->zi; hc.watsat = col->watsat; hc.bsw = col->bsw;
    hc.sucsat = col->sucsat; hc.hksat = col->hksat; hc.eff_porosity = col->eff_porosity; hc.nbedrock = col->nbedrock;
    hc.h2osoi_liq = col->h2osoi_liq; hc.h2osoi_ice = col->h2osoi_ice; hc.h2osfc = col->h2osfc;
    hc.icefrac = col->icefrac;
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    ctsm_col_state dcol{};
    ctsm_drainage_state ddr{};
    const int ncf = (int)(sizeof(kColFields) / sizeof(kColFields[0]));
    int rc = mirror(ctx, kColFields, ncf, &hc, &dcol, ctx->colmir, lay->begc_alloc, lay->ldc, cr, true);
    if (!rc) rc = mirror(ctx, kDr, kNDr, dr, &ddr, ctx->drm, lay->begc_alloc, lay->ldc, cr, true);
    if (rc) return rc;
    const int32_t *dfc = nullptr;
    if ((rc = upload_filter(ctx, n, filter, &dfc))) return rc;
    launch_drainage(ctx->stream, *lay, ctx->cfg, *prm, dcol, ddr, n, dfc, dtime, ctx->err);
    if ((rc = finish_launch(ctx))) return rc;
    if ((rc = mirror(ctx, kColFields, ncf, &hc, &dcol, ctx->colmir, lay->begc_alloc, lay->ldc, cr, false))) return rc;
    if ((rc = mirror(ctx, kDr, kNDr, dr, &ddr, ctx->drm, lay->begc_alloc, lay->ldc, cr, false))) return rc;
    return sync_and_check(ctx);
}

int ctsm_set_phs_params(ctsm_ctx *ctx, const ctsm_phs_params *prm)
{
    if (!ctx || !prm || prm->npft < 1 || prm->npft > CTSM_MAX_PFT) return CTSM_E_ARG;
    const double *rows[] = {prm->krmax, prm->root_radius, prm->root_density, prm->froot_leaf};
    for (const double *a : rows) if (!a) return CTSM_E_ARG;
    if (!prm->psi50 || !prm->ck) return CTSM_E_ARG;
    // rows: krmax, root_radius, root_density, froot_leaf, psi50, ck and kmax of the 4 segments, crop
    const int n = prm->npft;
    std::vector<double> tab((size_t)17 * CTSM_MAX_PFT, 0.0);
    for (int r = 0; r < 4; ++r)
        for (int i = 0; i < n; ++i) tab[(size_t)r * CTSM_MAX_PFT + i] = rows[r][i];
    for (int k = 0; k < 4; ++k)
        for (int i = 0; i < n; ++i) {
            tab[(size_t)(4 + k) * CTSM_MAX_PFT + i] = prm->psi50[(size_t)k * n + i];
            tab[(size_t)(8 + k) * CTSM_MAX_PFT + i] = prm->ck[(size_t)k * n + i];
            if (prm->kmax) tab[(size_t)(12 + k) * CTSM_MAX_PFT + i] = prm->kmax[(size_t)k * n + i];
        }
    if (prm->crop)
        for (int i = 0; i < n; ++i) tab[(size_t)16 * CTSM_MAX_PFT + i] = prm->crop[i];
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e);
    if ((e = ctx->phs_tab.ensure(tab.size() * sizeof(double))) != hipSuccess) return hip_fail(ctx, e);
    if ((e = hipMemcpyAsync(ctx->phs_tab.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice,
                            ctx->stream)) != hipSuccess)
        return hip_fail(ctx, e);
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return hip_fail(ctx, e);
    ctx->phs_npft = n;
    ctx->phs_kmax = prm->kmax != nullptr;
    return CTSM_OK;
}

// The members each PHS call reads or writes; the rest of a caller's struct is
// neither checked nor copied.
static ctsm_phs_state phs_pick(const ctsm_phs_state *s, bool conductance)
{
    ctsm_phs_state o{};
    o.z = s->z; o.smp_l = s->smp_l; o.k_soil_root = s->k_soil_root;
    if (conductance) {
        o.itype = s->itype; o.column = s->column; o.froot_carbon = s->froot_carbon; o.tsai = s->tsai;
        o.tlai = s->tlai; o.rootfr = s->rootfr; o.root_conductance = s->root_conductance;
        o.soil_conductance = s->soil_conductance; o.dz = s->dz; o.hksat = s->hksat; o.hk_l = s->hk_l;
    } else {
        o.active = s->active; o.frac_veg_nosno = s->frac_veg_nosno; o.wtcol = s->wtcol; o.vegwp = s->vegwp;
        o.qflx_hydr_redist = s->qflx_hydr_redist; o.patchi = s->patchi; o.patchf = s->patchf;
        o.qflx_rootsoi = s->qflx_rootsoi; o.qflx_phs_neg = s->qflx_phs_neg;
    }
    return o;
}

static bool phs_complete(const ctsm_phs_state *s, bool conductance)
{
    ctsm_phs_state probe;
    std::memset(&probe, 0xff, sizeof(probe));
    const ctsm_phs_state need = phs_pick(&probe, conductance);   // non-null where the call uses the member
    for (size_t off = 0; off < sizeof(ctsm_phs_state); off += sizeof(void *))
        if (*(void *const *)((const char *)&need + off) && !*(void *const *)((const char *)s + off)) return false;
    return true;
}

static int phs_check(const ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *s, int32_t n,
                     bool conductance, bool has_filter)
{
    if (!ctx || !s || n < 0 || check_layout(lay) || lay->ldp <= 0) return CTSM_E_ARG;
    if (!has_filter && n > (conductance ? lay->ldp : lay->ldc)) return CTSM_E_ARG;
    if (conductance && ctx->phs_npft == 0) return CTSM_E_ARG;
    if (ctx->cfg.nlevsoi > MAX_NLEVSOI) return CTSM_E_ARG;
    return phs_complete(s, conductance) ? CTSM_OK : CTSM_E_ARG;
}

int ctsm_phs_soil_root_conductance_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *phs,
                                       int32_t n, const int32_t *filter_dev)
{
    int rc = phs_check(ctx, lay, phs, n, true, filter_dev != nullptr);
    if (rc) return rc;
    launch_phs_conductance(ctx->stream, *lay, ctx->cfg, *phs, (const double *)ctx->phs_tab.p, ctx->phs_npft, n,
                           filter_dev, ctx->err);
    return finish_launch(ctx);
}

static int canopy_phs_check(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                            const ctsm_canopy_state *can, const ctsm_phs_state *phs,
                            const ctsm_phs_canopy_state *phc, int32_t n, bool has_filter)
{
    if (!ctx || !col || !can || !phc || n < 0 || check_canopy_layout(ctx, lay) || (!has_filter && n > lay->ldp))
        return CTSM_E_ARG;
    if (!ctx->cfg.use_hydrstress) return CTSM_E_CONFIG;
    int rc = phs_check(ctx, lay, phs, n, true, has_filter);
    if (rc) return rc;
    if (!ctx->phs_kmax) return CTSM_E_ARG;   // calcstress needs kmax
    for (size_t off = 0; off < sizeof(ctsm_phs_canopy_state); off += sizeof(void *))
        if (!*(void *const *)((const char *)phc + off)) return CTSM_E_ARG;
    return CTSM_OK;
}

int ctsm_canopyfluxes_phs_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                              const ctsm_canopy_state *can, const ctsm_phs_state *phs,
                              const ctsm_phs_canopy_state *phc, int32_t n, const int32_t *filterp_dev, double dtime)
{
    int rc = canopy_phs_check(ctx, lay, col, can, phs, phc, n, filterp_dev != nullptr);
    if (rc) return rc;
    // the conductance loop at the head of PhotosynthesisHydraulicStress, then the iteration
    launch_phs_conductance(ctx->stream, *lay, ctx->cfg, *phs, (const double *)ctx->phs_tab.p, ctx->phs_npft, n,
                           filterp_dev, ctx->err);
    rc = run_canopy(ctx, lay, *col, *can, n, filterp_dev, dtime, false, nullptr, 0, false, phs, phc);
    if (rc) return rc;
    return finish_launch(ctx);
}

int ctsm_phs_root_sink_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *phs, int32_t n,
                           const int32_t *filter_dev)
{
    int rc = phs_check(ctx, lay, phs, n, false, filter_dev != nullptr);
    if (rc) return rc;
    launch_phs_root_sink(ctx->stream, *lay, ctx->cfg, *phs, n, filter_dev, ctx->err);
    return finish_launch(ctx);
}

// Drop-in form of both: mirror the members the call uses over the whole
// allocation, run, copy the outputs back.
static int phs_dropin(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *phs, int32_t n,
                      const int32_t *filter, bool conductance)
{
    int rc = phs_check(ctx, lay, phs, n, conductance, filter != nullptr);
    if (rc) return rc;
    const int beg = conductance ? lay->begp_alloc : lay->begc_alloc, ld = conductance ? lay->ldp : lay->ldc;
    for (int i = 0; filter && i < n; ++i)
        if (filter[i] < beg || filter[i] >= beg + ld) return CTSM_E_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    const ctsm_phs_state h = phs_pick(phs, conductance);
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    const Range pr{lay->begp_alloc, lay->begp_alloc + lay->ldp - 1};
    ctsm_phs_state d{};
    rc = mirror(ctx, kPhsP, kNPhsP, &h, &d, ctx->phsp, lay->begp_alloc, lay->ldp, pr, true);
    if (!rc) rc = mirror(ctx, kPhsC, kNPhsC, &h, &d, ctx->phsc, lay->begc_alloc, lay->ldc, cr, true);
    if (rc) return rc;
    const int32_t *df = nullptr;
    if ((rc = upload_filter(ctx, n, filter, &df))) return rc;
    if (conductance)
        launch_phs_conductance(ctx->stream, *lay, ctx->cfg, d, (const double *)ctx->phs_tab.p, ctx->phs_npft, n, df,
                               ctx->err);
    else
        launch_phs_root_sink(ctx->stream, *lay, ctx->cfg, d, n, df, ctx->err);
    if ((rc = finish_launch(ctx))) return rc;
    if ((rc = mirror(ctx, kPhsP, kNPhsP, &h, &d, ctx->phsp, lay->begp_alloc, lay->ldp, pr, false))) return rc;
    if ((rc = mirror(ctx, kPhsC, kNPhsC, &h, &d, ctx->phsc, lay->begc_alloc, lay->ldc, cr, false))) return rc;
    return sync_and_check(ctx);
}

int ctsm_phs_soil_root_conductance(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *phs, int32_t n,
                                   const int32_t *filterp)
{
    return phs_dropin(ctx, lay, phs, n, filterp, true);
}

int ctsm_phs_root_sink(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_state *phs, int32_t n,
                       const int32_t *filterc)
{
    return phs_dropin(ctx, lay, phs, n, filterc, false);
}

// Drop-in form of ctsm_canopyfluxes_phs_dev: the column, canopy, conductance
// and PHS canopy members mirrored over the whole allocation, the writable ones
// copied back.  Arrays two structs share (vegwp, the aliased inputs) get one
// mirror each; only one of them is writable, so the copy back is unambiguous.
int ctsm_canopyfluxes_phs(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                          const ctsm_canopy_state *can, const ctsm_phs_state *phs, const ctsm_phs_canopy_state *phc,
                          int32_t n, const int32_t *filterp, double dtime)
{
    int rc = canopy_phs_check(ctx, lay, col, can, phs, phc, n, filterp != nullptr);
    if (rc) return rc;
    for (int i = 0; filterp && i < n; ++i)
        if (filterp[i] < lay->begp_alloc || filterp[i] >= lay->begp_alloc + lay->ldp) return CTSM_E_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    const Range pr{lay->begp_alloc, lay->begp_alloc + lay->ldp - 1};
    const ctsm_phs_state h = phs_pick(phs, true);
    ctsm_col_state dcol{};
    ctsm_canopy_state dcan{};
    ctsm_phs_state dphs{};
    ctsm_phs_canopy_state dphc{};
    const int ncolf = (int)(sizeof(kColFields) / sizeof(kColFields[0]));
    auto move = [&](bool to_device) {
        int r = mirror(ctx, kColFields, ncolf, col, &dcol, ctx->colmir, lay->begc_alloc, lay->ldc, cr, to_device);
        if (!r) r = canopy_mirror(ctx, lay, can, &dcan, to_device);
        if (!r) r = mirror(ctx, kPhsP, kNPhsP, &h, &dphs, ctx->phsp, lay->begp_alloc, lay->ldp, pr, to_device);
        if (!r) r = mirror(ctx, kPhsC, kNPhsC, &h, &dphs, ctx->phsc, lay->begc_alloc, lay->ldc, cr, to_device);
        if (!r) r = mirror(ctx, kPhc, kNPhc, phc, &dphc, ctx->phcm, lay->begp_alloc, lay->ldp, pr, to_device);
        return r;
    };
    if ((rc = move(true))) return rc;
    const int32_t *dfilt;
    if ((rc = upload_filter(ctx, n, filterp, &dfilt))) return rc;
    launch_phs_conductance(ctx->stream, *lay, ctx->cfg, dphs, (const double *)ctx->phs_tab.p, ctx->phs_npft, n, dfilt,
                           ctx->err);
    if ((rc = run_canopy(ctx, lay, dcol, dcan, n, dfilt, dtime, false, nullptr, 0, false, &dphs, &dphc))) return rc;
    if ((rc = finish_launch(ctx))) return rc;
    if ((rc = move(false))) return rc;
    return sync_and_check(ctx);
}

static int stress_check(const ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_stress_state *st, int32_t n,
                        bool has_filter)
{
    if (!ctx || !st || n < 0 || check_layout(lay) || lay->ldp <= 0) return CTSM_E_ARG;
    if ((!has_filter && n > lay->ldp) || ctx->phs_npft == 0 || !ctx->phs_kmax) return CTSM_E_ARG;
    if (ctx->cfg.nlevsoi > MAX_NLEVSOI) return CTSM_E_ARG;
    for (size_t off = 0; off < sizeof(ctsm_phs_stress_state); off += sizeof(void *))
        if (!*(void *const *)((const char *)st + off)) return CTSM_E_ARG;
    return CTSM_OK;
}

int ctsm_phs_calcstress_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_stress_state *st, int32_t n,
                            const int32_t *filter_dev)
{
    int rc = stress_check(ctx, lay, st, n, filter_dev != nullptr);
    if (rc) return rc;
    launch_phs_calcstress(ctx->stream, *lay, ctx->cfg, *st, (const double *)ctx->phs_tab.p, ctx->phs_npft, n,
                          filter_dev, ctx->err);
    return finish_launch(ctx);
}

int ctsm_phs_calcstress(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_phs_stress_state *st, int32_t n,
                        const int32_t *filter)
{
    int rc = stress_check(ctx, lay, st, n, filter != nullptr);
    if (rc) return rc;
    for (int i = 0; filter && i < n; ++i)
        if (filter[i] < lay->begp_alloc || filter[i] >= lay->begp_alloc + lay->ldp) return CTSM_E_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    const Range pr{lay->begp_alloc, lay->begp_alloc + lay->ldp - 1};
    ctsm_phs_stress_state d{};
    rc = mirror(ctx, kStP, kNStP, st, &d, ctx->stp, lay->begp_alloc, lay->ldp, pr, true);
    if (!rc) rc = mirror(ctx, kStC, kNStC, st, &d, ctx->stc, lay->begc_alloc, lay->ldc, cr, true);
    if (rc) return rc;
    const int32_t *df = nullptr;
    if ((rc = upload_filter(ctx, n, filter, &df))) return rc;
    launch_phs_calcstress(ctx->stream, *lay, ctx->cfg, d, (const double *)ctx->phs_tab.p, ctx->phs_npft, n, df,
                          ctx->err);
    if ((rc = finish_launch(ctx))) return rc;
    if ((rc = mirror(ctx, kStP, kNStP, st, &d, ctx->stp, lay->begp_alloc, lay->ldp, pr, false))) return rc;
    return sync_and_check(ctx);
}

static int land_step_full(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                          const ctsm_patch_state *pch, const ctsm_canopy_state *can, const ctsm_soilflux_state *sf,
                          const ctsm_surface_water_state *sw, const ctsm_hydro_params *hprm,
                          const ctsm_water_table_state *wt, const ctsm_drainage_state *dr, int32_t nc,
                          const int32_t *filterc_dev, int32_t np, const int32_t *filterp_dev, double dtime,
                          bool bounds_are_filter)
{
    if (!ctx || !col || !pch || !can || !sf || nc < 0 || np < 0 || check_canopy_layout(ctx, lay) ||
        (!filterc_dev && nc > lay->ldc) || (!filterp_dev && np > lay->ldp) || !sf_complete(sf))
        return CTSM_E_ARG;
    if (sw && (!hprm || !sw_complete(sw) || sw->qflx_infl != col->qflx_infl || sw->h2osfc != col->h2osfc))
        return CTSM_E_ARG;
    if ((wt || dr) && ctx->cfg.lower_boundary_condition == CTSM_BC_WATERTABLE) return CTSM_E_CONFIG;
    if (wt && (!wt_complete(wt) || (sw && (sw->zwt != wt->zwt || sw->zwt_perched != wt->zwt_perched ||
                                           sw->frost_table != wt->frost_table))))
        return CTSM_E_ARG;
    if (dr && (!hprm || !dr_complete(dr) || (wt && (dr->zwt != wt->zwt || dr->zwt_perched != wt->zwt_perched ||
                                                    dr->frost_table != wt->frost_table)) ||
               (sw && (dr->qflx_surf != sw->qflx_surf || dr->topo_slope != sw->topo_slope))))
        return CTSM_E_ARG;
    // the arrays one routine writes and the next reads must coincide
    if (pch->dlrad != can->dlrad || pch->cgrnd != can->cgrnd || pch->eflx_sh_grnd != can->eflx_sh_grnd ||
        pch->eflx_sh_snow != can->eflx_sh_snow || pch->eflx_sh_soil != can->eflx_sh_soil ||
        pch->eflx_sh_h2osfc != can->eflx_sh_h2osfc || pch->qflx_evap_soi != can->qflx_evap_soi ||
        pch->qflx_ev_snow != can->qflx_ev_snow || pch->qflx_ev_soil != can->qflx_ev_soil ||
        pch->qflx_ev_h2osfc != can->qflx_ev_h2osfc || col->qflx_rootsoi != can->qflx_rootsoi_col ||
        sf->eflx_sh_grnd != can->eflx_sh_grnd || sf->qflx_evap_soi != can->qflx_evap_soi ||
        sf->qflx_ev_snow != can->qflx_ev_snow || sf->qflx_ev_soil != can->qflx_ev_soil ||
        sf->qflx_ev_h2osfc != can->qflx_ev_h2osfc || sf->cgrnds != can->cgrnds || sf->cgrndl != can->cgrndl ||
        sf->ulrad != can->ulrad || sf->eflx_sh_veg != can->eflx_sh_veg || sf->eflx_sh_stem != can->eflx_sh_stem ||
        sf->qflx_evap_veg != can->qflx_evap_veg || sf->qflx_tran_veg != can->qflx_tran_veg ||
        sf->t_skin != can->t_skin || sf->active != can->active)
        return CTSM_E_ARG;
    int rc;
    const bool plain = ctx->tun.plain_routines != 0;
    if ((rc = ensure_work(ctx, land_work_doubles(ctx, nc)))) return rc;
    launch_save_before(ctx->stream, *lay, ctx->cfg, *col, *sf, nc, filterc_dev);
    if ((rc = run_canopy(ctx, lay, *col, *can, np, filterp_dev, dtime, true, nullptr, 0, bounds_are_filter)))
        return rc;
    const View V = make_view(ctx, *lay);
    if (plain) {
        launch_soiltemp_plain(ctx->stream, V, ctx->prm, *col, *pch, nc, filterc_dev, dtime, (double *)ctx->work.p,
                              nc, ctx->err);
        launch_soilwater_fractions(ctx->stream, V, *col, nc, filterc_dev);
    } else {
        launch_soiltemp_fractions(ctx->stream, V, ctx->prm, *col, *pch, nc, filterc_dev, dtime,
                                  (double *)ctx->work.p, nc, ctx->err, ctx->tun.soiltemp_block);
    }
    launch_soilfluxes(ctx->stream, *lay, ctx->cfg, *col, *pch, *sf, nc, filterc_dev, np, filterp_dev, dtime, ctx->err);
    launch_root_sink(ctx->stream, *lay, ctx->cfg, *col, *can, nc, filterc_dev, ctx->err);
    if (sw) {
        launch_patch2col3(ctx->stream, *lay, *col, can->active, can->wtcol, sf->qflx_liqevap_from_top_layer,
                          can->qflx_ev_soil, can->qflx_ev_h2osfc, sw->qflx_liqevap_from_top_layer, sw->qflx_ev_soil,
                          sw->qflx_ev_h2osfc, nc, filterc_dev, ctx->err);
        launch_surface_water(ctx->stream, *lay, ctx->cfg.e_ice, *hprm, *col, *sw, nc, filterc_dev, dtime, ctx->err);
    }
    if (plain) {
        launch_soilwater_plain(ctx->stream, V, ctx->prm, *col, nc, filterc_dev, dtime, (double *)ctx->work.p, nc,
                               ctx->err);
    } else {
        // columns sorted by their previous sub-step count, so a wave's columns
        // sub-step alike (the adaptive group size keeps the filter's depth order)
        const int32_t *rf = filterc_dev;
        rf = launch_richards_order(ctx->stream, V, *col, nc, filterc_dev, ctx->work.p, ctx->tun.richards_lanes);
        launch_soilwater(ctx->stream, V, ctx->prm, *col, nc, rf, dtime, ctx->err, ctx->tun.richards_lanes,
                         ctx->tun.richards_block);
    }
    if (wt) {
        launch_patch2col3(ctx->stream, *lay, *col, can->active, can->wtcol, sf->qflx_liqdew_to_top_layer,
                          sf->qflx_soliddew_to_top_layer, sf->qflx_solidevap_from_top_layer,
                          wt->qflx_liqdew_to_top_layer, wt->qflx_soliddew_to_top_layer,
                          wt->qflx_solidevap_from_top_layer, nc, filterc_dev, ctx->err);
        launch_water_table(ctx->stream, *lay, ctx->cfg, *col, *wt, nc, filterc_dev, dtime, ctx->err);
    }
    if (dr) launch_drainage(ctx->stream, *lay, ctx->cfg, *hprm, *col, *dr, nc, filterc_dev, dtime, ctx->err);
    return finish_launch(ctx);
}

int ctsm_land_step_full_dev(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                            const ctsm_patch_state *pch, const ctsm_canopy_state *can, const ctsm_soilflux_state *sf,
                            const ctsm_surface_water_state *sw, const ctsm_hydro_params *hprm,
                            const ctsm_water_table_state *wt, const ctsm_drainage_state *dr, int32_t nc,
                            const int32_t *filterc_dev, int32_t np, const int32_t *filterp_dev, double dtime)
{
    return land_step_full(ctx, lay, col, pch, can, sf, sw, hprm, wt, dr, nc, filterc_dev, np, filterp_dev, dtime,
                          false);
}

int ctsm_canopyfluxes(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col,
                      const ctsm_canopy_state *can, int32_t n, const int32_t *filterp, double dtime)
{
    if (!ctx || !col || !can || n < 0 || check_canopy_layout(ctx, lay) || (!filterp && n > lay->ldp))
        return CTSM_E_ARG;
    if (ctx->cfg.use_hydrstress) return CTSM_E_CONFIG;   // before any copy: the PHS entries serve this context
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    for (int i = 0; filterp && i < n; ++i)
        if (filterp[i] < lay->begp_alloc || filterp[i] >= lay->begp_alloc + lay->ldp) return CTSM_E_ARG;
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    ctsm_col_state dcol{};
    ctsm_canopy_state dcan{};
    int rc = mirror(ctx, kColFields, (int)(sizeof(kColFields) / sizeof(kColFields[0])), col, &dcol, ctx->colmir,
                    lay->begc_alloc, lay->ldc, cr, true);
    if (rc) return rc;
    if ((rc = canopy_mirror(ctx, lay, can, &dcan, true))) return rc;
    const int32_t *dfilt;
    if ((rc = upload_filter(ctx, n, filterp, &dfilt))) return rc;
    if ((rc = run_canopy(ctx, lay, dcol, dcan, n, dfilt, dtime, false))) return rc;
    if ((rc = finish_launch(ctx))) return rc;
    rc = mirror(ctx, kColFields, (int)(sizeof(kColFields) / sizeof(kColFields[0])), col, &dcol, ctx->colmir,
                lay->begc_alloc, lay->ldc, cr, false);
    if (rc) return rc;
    if ((rc = canopy_mirror(ctx, lay, can, &dcan, false))) return rc;
    return sync_and_check(ctx);
}

int ctsm_root_sink(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col, const ctsm_canopy_state *can,
                   int32_t n, const int32_t *filterc)
{
    if (!ctx || !col || !can || n < 0 || check_canopy_layout(ctx, lay) || (!filterc && n > lay->ldc))
        return CTSM_E_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    for (int i = 0; filterc && i < n; ++i)
        if (filterc[i] < lay->begc_alloc || filterc[i] >= lay->begc_alloc + lay->ldc) return CTSM_E_ARG;
    const Range cr{lay->begc_alloc, lay->begc_alloc + lay->ldc - 1};
    ctsm_col_state dcol{};
    ctsm_canopy_state dcan{};
    int rc = mirror(ctx, kColFields, (int)(sizeof(kColFields) / sizeof(kColFields[0])), col, &dcol, ctx->colmir,
                    lay->begc_alloc, lay->ldc, cr, true);
    if (rc) return rc;
    if ((rc = canopy_mirror(ctx, lay, can, &dcan, true))) return rc;
    const int32_t *dfilt;
    if ((rc = upload_filter(ctx, n, filterc, &dfilt))) return rc;
    launch_root_sink(ctx->stream, *lay, ctx->cfg, dcol, dcan, n, dfilt, ctx->err);
    if ((rc = finish_launch(ctx))) return rc;
    if ((rc = canopy_mirror(ctx, lay, can, &dcan, false))) return rc;
    return sync_and_check(ctx);
}

int ctsm_dev_alloc(ctsm_ctx *ctx, uint64_t bytes, void **out)
{
    if (!ctx || !out) return CTSM_E_ARG;
    *out = nullptr;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMalloc(out, std::max<uint64_t>(bytes, 1));
    return e == hipSuccess ? CTSM_OK : hip_fail(ctx, e);
}

int ctsm_dev_free(ctsm_ctx *ctx, void *ptr)
{
    if (!ctx) return CTSM_E_ARG;
    if (!ptr) return CTSM_OK;
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipFree(ptr);
    return e == hipSuccess ? CTSM_OK : hip_fail(ctx, e);
}

int ctsm_memcpy(ctsm_ctx *ctx, void *dst, const void *src, uint64_t bytes, int32_t kind)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return CTSM_E_ARG;
    hipMemcpyKind k;
    switch (kind) {
    case 1: k = hipMemcpyHostToDevice; break;
    case 2: k = hipMemcpyDeviceToHost; break;
    case 3: k = hipMemcpyDeviceToDevice; break;
    default: return CTSM_E_ARG;
    }
    if (bytes == 0) return CTSM_OK;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, k, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? CTSM_OK : hip_fail(ctx, e);
}

int ctsm_memcpy2d(ctsm_ctx *ctx, void *dst, uint64_t dpitch, const void *src, uint64_t spitch, uint64_t width,
                  uint64_t height, int32_t kind)
{
    if (!ctx || (!dst && width * height) || (!src && width * height) || width > dpitch || width > spitch)
        return CTSM_E_ARG;
    if (width == 0 || height == 0) return CTSM_OK;
    hipMemcpyKind k;
    switch (kind) {
    case 1: k = hipMemcpyHostToDevice; break;
    case 2: k = hipMemcpyDeviceToHost; break;
    case 3: k = hipMemcpyDeviceToDevice; break;
    default: return CTSM_E_ARG;
    }
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e == hipSuccess) e = hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, k, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? CTSM_OK : hip_fail(ctx, e);
}

int ctsm_dev_copy(ctsm_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return CTSM_E_ARG;
    if (bytes == 0) return CTSM_OK;
    const bool aligned = (((uintptr_t)dst | (uintptr_t)src | (uintptr_t)bytes) & 15) == 0;
    if (aligned) {
        const size_t n16 = bytes / 16;
        const size_t blocks = std::min<size_t>((n16 + 255) / 256, 256 * 8 * 4);
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (uint4 *)dst,
                           (const uint4 *)src, n16);
    } else {
        const size_t blocks = std::min<size_t>((bytes + 255) / 256, 256 * 8 * 4);
        hipLaunchKernelGGL(k_copy1, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (char *)dst,
                           (const char *)src, (size_t)bytes);
    }
    return finish_launch(ctx);
}

int ctsm_event_record(ctsm_ctx *ctx, void **event)
{
    if (!ctx || !event) return CTSM_E_ARG;
    hipEvent_t ev = (hipEvent_t)*event;
    hipError_t e = hipSuccess;
    if (!ev) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipEventRecord(ev, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e);
    *event = (void *)ev;
    return CTSM_OK;
}

int ctsm_event_elapsed(void *ev_start, void *ev_stop, float *ms)
{
    if (!ev_start || !ev_stop || !ms) return CTSM_E_ARG;
    if (hipEventSynchronize((hipEvent_t)ev_stop) != hipSuccess) return CTSM_E_HIP;
    return hipEventElapsedTime(ms, (hipEvent_t)ev_start, (hipEvent_t)ev_stop) == hipSuccess ? CTSM_OK : CTSM_E_HIP;
}

int ctsm_event_destroy(void *event)
{
    if (!event) return CTSM_OK;
    return hipEventDestroy((hipEvent_t)event) == hipSuccess ? CTSM_OK : CTSM_E_HIP;
}

int ctsm_sync(ctsm_ctx *ctx)
{
    if (!ctx) return CTSM_E_ARG;
    return sync_and_check(ctx);
}

void *ctsm_stream(ctsm_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

}  // extern "C"

// ---- history buffers ---------------------------------------------------------

static int hist_call(ctsm_ctx *ctx, const ctsm_hist_field *fields, int32_t n, int op)
{
    if (!ctx || n < 0 || (n > 0 && !fields)) return CTSM_E_ARG;
    for (int32_t i = 0; i < n; ++i) {
        const ctsm_hist_field &f = fields[i];
        const int a = f.avgflag;
        if (f.n1d < 0 || f.num2d < 0 || (long long)f.n1d * f.num2d > INT32_MAX ||
            (a != 'A' && a != 'S' && a != 'I' && a != 'X' && a != 'M'))
            return CTSM_E_ARG;
        if ((long long)f.n1d * f.num2d == 0) continue;
        if (!f.hbuf || !f.nacs || (op == 0 && !f.field) || (f.num2d > 1 && f.ld < f.n1d)) return CTSM_E_ARG;
    }
    launch_hist(ctx->stream, fields, n, op);
    return finish_launch(ctx);
}

int ctsm_hist_update_dev(ctsm_ctx *ctx, const ctsm_hist_field *fields, int32_t n) { return hist_call(ctx, fields, n, 0); }
int ctsm_hist_normalize_dev(ctsm_ctx *ctx, const ctsm_hist_field *fields, int32_t n) { return hist_call(ctx, fields, n, 1); }
int ctsm_hist_zero_dev(ctsm_ctx *ctx, const ctsm_hist_field *fields, int32_t n) { return hist_call(ctx, fields, n, 2); }

// ---- checkpoints ---------------------------------------------------------------

static int ckpt_call(ctsm_ctx *ctx, const char *path, const ctsm_ckpt_field *fields, int32_t n, bool write)
{
    if (!ctx || !path || n < 0 || (n > 0 && !fields)) return CTSM_E_ARG;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    // the state a checkpoint holds is the one after every enqueued step; a
    // failed step is reported instead of written
    int rc = sync_and_check(ctx);
    if (rc) return rc;
    rc = write ? ckpt_write(ctx->stream, path, fields, n, &he) : ckpt_read(ctx->stream, path, fields, n, &he);
    if (rc == CTSM_E_HIP) return hip_fail(ctx, he);
    ctx->last_code = rc;
    ctx->last_index = 0;
    return rc;
}

int ctsm_checkpoint_write(ctsm_ctx *ctx, const char *path, const ctsm_ckpt_field *fields, int32_t n)
{
    return ckpt_call(ctx, path, fields, n, true);
}

int ctsm_checkpoint_read(ctsm_ctx *ctx, const char *path, const ctsm_ckpt_field *fields, int32_t n)
{
    return ckpt_call(ctx, path, fields, n, false);
}

// ---- PFT rows of one member ------------------------------------------------------

int ctsm_set_pft_rows(ctsm_ctx *ctx, int32_t first_row, const ctsm_pft_params *pft)
{
    if (!ctx || !pft || !ctx->have_canopy || first_row < 0 || pft->npft < 1 ||
        first_row + pft->npft > CTSM_MAX_PFT)
        return CTSM_E_ARG;
    const double *cols[] = {pft->dleaf, pft->c3psn, pft->slatop, pft->leafcn, pft->flnr, pft->fnitr, pft->mbbopt,
                            pft->medlynslope, pft->medlynintercept, pft->smpso, pft->smpsc, pft->theta_cj,
                            pft->z0v_Cr, pft->z0v_Cs, pft->z0v_c, pft->z0v_cw, pft->z0v_LAImax};
    const double *cols2[] = {pft->dbh, pft->fbw, pft->nstem, pft->rstem_per_dbh, pft->wood_density};
    for (const double *a : cols) if (!a) return CTSM_E_ARG;
    for (const double *a : cols2) if (!a) return CTSM_E_ARG;
    if (!pft->is_tree || !pft->is_shrub) return CTSM_E_ARG;
    const int n = pft->npft;
    std::vector<double> rows((size_t)24 * n);
    for (int f = 0; f < 17; ++f) std::memcpy(&rows[(size_t)f * n], cols[f], n * sizeof(double));
    for (int i = 0; i < n; ++i) {
        rows[(size_t)17 * n + i] = (double)pft->is_tree[i];
        rows[(size_t)18 * n + i] = (double)pft->is_shrub[i];
    }
    for (int f = 0; f < 5; ++f) std::memcpy(&rows[(size_t)(19 + f) * n], cols2[f], n * sizeof(double));
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e);
    // one strided copy: field f's rows go to table row f, columns first_row..
    e = hipMemcpy2DAsync((double *)ctx->pft.p + first_row, CTSM_MAX_PFT * sizeof(double), rows.data(),
                         n * sizeof(double), n * sizeof(double), 24, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e);
    return CTSM_OK;
}

// ---- the ensemble server: continuous batching in a resident pool ---------------

namespace {

// First-fit allocator of index ranges [0, cap).
struct RangeSet {
    std::vector<std::pair<int64_t, int64_t>> free;     // (start, length), sorted
    void init(int64_t cap)
    {
        free.clear();
        if (cap > 0) free.push_back({0, cap});
    }
    int64_t alloc(int64_t n)
    {
        if (n <= 0) return 0;
        for (size_t i = 0; i < free.size(); ++i)
            if (free[i].second >= n) {
                const int64_t s = free[i].first;
                if (free[i].second == n) free.erase(free.begin() + i);
                else free[i] = {s + n, free[i].second - n};
                return s;
            }
        return -1;
    }
    bool take(int64_t s, int64_t n)
    {
        if (n <= 0) return true;
        for (size_t i = 0; i < free.size(); ++i) {
            const int64_t a = free[i].first, l = free[i].second;
            if (a <= s && s + n <= a + l) {
                std::vector<std::pair<int64_t, int64_t>> parts;
                if (s > a) parts.push_back({a, s - a});
                if (a + l > s + n) parts.push_back({s + n, a + l - s - n});
                free.erase(free.begin() + i);
                free.insert(free.begin() + i, parts.begin(), parts.end());
                return true;
            }
        }
        return false;
    }
    void release(int64_t s, int64_t n)
    {
        if (n <= 0) return;
        auto it = std::lower_bound(free.begin(), free.end(), std::make_pair(s, (int64_t)0));
        free.insert(it, {s, n});
        std::vector<std::pair<int64_t, int64_t>> m;
        for (auto &r : free) {
            if (!m.empty() && m.back().first + m.back().second == r.first) m.back().second += r.second;
            else m.push_back(r);
        }
        free.swap(m);
    }
};

struct Member {
    int32_t ncol = 0, npatch = 0, npft = 0, nsteps = 0, left = 0;
    int32_t col0 = -1, pat0 = -1, pft0 = -1;
    int32_t state = CTSM_SERVE_QUEUED;
    int64_t finish_tick = -1;
};

constexpr int kRing = 3;          // filter buffers (at most kMaxInflight ticks use them at once)
constexpr int kMaxInflight = 2;   // ticks enqueued and not yet retired
// balance words per step: |errsoi| max and its column, |errseb| max and its
// patch, |errlon| max and its patch
constexpr int kBalWords = 6;

struct Tick {
    int64_t id = 0;
    int slot = 0;                 // balance / error result slot (= ring slot of its filter)
    int steps = 0;
    hipEvent_t ev = nullptr;
};

}  // namespace

struct ctsm_server {
    ctsm_ctx *ctx = nullptr;
    ctsm_layout lay{};
    ctsm_col_state col{};
    ctsm_patch_state pch{};
    ctsm_canopy_state can{};
    ctsm_soilflux_state sf{};
    ctsm_surface_water_state sw{};
    ctsm_water_table_state wt{};
    ctsm_drainage_state dr{};
    ctsm_hydro_params hprm{};
    bool have_sw = false, have_wt = false, have_dr = false;
    bool check_energy = false;        // the state holds what errseb / errlon read
    double dtime = 1800.0;
    int32_t max_steps = 8;
    RangeSet cols, pats, pfts;
    int32_t pft_first = 0;        // first PFT row members may take
    std::map<int64_t, Member> members;
    std::deque<int64_t> queue;
    int64_t next_ticket = 1, next_tick = 0;
    // filters: device buffers and pinned host staging, one per ring slot
    int32_t *fc_dev[kRing] = {}, *fp_dev[kRing] = {}, *fc_host[kRing] = {}, *fp_host[kRing] = {};
    hipEvent_t ring_ev[kRing] = {};   // the last tick that read the slot's filter
    bool ring_used[kRing] = {};
    int cur = -1;
    int32_t nc = 0, np = 0;
    bool dirty = true;
    // per tick slot: max_steps x kBalWords balance words and the error word
    unsigned long long *bal_dev = nullptr, *bal_host = nullptr;
    ErrWord *err_host = nullptr;
    std::deque<Tick> inflight;
    double worst_errsoi = 0.0;
    int32_t worst_col = 0;
    double worst_errseb = 0.0, worst_errlon = 0.0;
    int32_t worst_seb_patch = 0, worst_lon_patch = 0;
    int64_t column_steps = 0;
    int32_t err_code = 0, err_index = 0;
};

static void serve_free(ctsm_server *s)
{
    for (int k = 0; k < kRing; ++k) {
        if (s->fc_dev[k]) (void)hipFree(s->fc_dev[k]);
        if (s->fp_dev[k]) (void)hipFree(s->fp_dev[k]);
        if (s->fc_host[k]) (void)hipHostFree(s->fc_host[k]);
        if (s->fp_host[k]) (void)hipHostFree(s->fp_host[k]);
        if (s->ring_ev[k]) (void)hipEventDestroy(s->ring_ev[k]);
    }
    for (auto &t : s->inflight)
        if (t.ev) (void)hipEventDestroy(t.ev);
    if (s->bal_dev) (void)hipFree(s->bal_dev);
    if (s->bal_host) (void)hipHostFree(s->bal_host);
    if (s->err_host) (void)hipHostFree(s->err_host);
    delete s;
}

int ctsm_serve_create(ctsm_ctx *ctx, const ctsm_layout *lay, const ctsm_col_state *col, const ctsm_patch_state *pch,
                      const ctsm_canopy_state *can, const ctsm_soilflux_state *sf, const ctsm_surface_water_state *sw,
                      const ctsm_hydro_params *hprm, const ctsm_water_table_state *wt, const ctsm_drainage_state *dr,
                      int32_t pft_first, int32_t max_steps_per_tick, double dtime, ctsm_server **out)
{
    if (!ctx || !lay || !col || !pch || !can || !sf || !out || check_canopy_layout(ctx, lay) || !sf_complete(sf) ||
        max_steps_per_tick < 1 || max_steps_per_tick > 64 || !(dtime > 0.0) || pft_first < 0 ||
        pft_first >= CTSM_MAX_PFT || ((sw || dr) && !hprm))
        return CTSM_E_ARG;
    if ((wt || dr) && ctx->cfg.lower_boundary_condition == CTSM_BC_WATERTABLE) return CTSM_E_CONFIG;
    *out = nullptr;
    std::lock_guard<std::mutex> g(ctx->mu);
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    ctsm_server *s = new ctsm_server;
    s->ctx = ctx;
    s->lay = *lay;
    s->col = *col;
    s->pch = *pch;
    s->can = *can;
    s->sf = *sf;
    if (sw) { s->sw = *sw; s->have_sw = true; }
    if (wt) { s->wt = *wt; s->have_wt = true; }
    if (dr) { s->dr = *dr; s->have_dr = true; }
    if (hprm) s->hprm = *hprm;
    s->check_energy = col->forc_lwrad && pch->sabg_chk && can->sabv && can->dhsdt_canopy && can->active &&
                      can->column && sf->eflx_lwrad_out && sf->eflx_lwrad_net && sf->eflx_sh_tot &&
                      sf->eflx_lh_tot && sf->eflx_soil_grnd;
    s->dtime = dtime;
    s->max_steps = max_steps_per_tick;
    s->cols.init(lay->ldc);
    s->pats.init(lay->ldp);
    s->pft_first = pft_first;
    s->pfts.init(CTSM_MAX_PFT - pft_first);
    bool ok = true;
    for (int k = 0; ok && k < kRing; ++k)
        ok = hipMalloc((void **)&s->fc_dev[k], (size_t)lay->ldc * sizeof(int32_t)) == hipSuccess &&
             hipMalloc((void **)&s->fp_dev[k], (size_t)lay->ldp * sizeof(int32_t)) == hipSuccess &&
             hipHostMalloc((void **)&s->fc_host[k], (size_t)lay->ldc * sizeof(int32_t), hipHostMallocDefault) ==
                 hipSuccess &&
             hipHostMalloc((void **)&s->fp_host[k], (size_t)lay->ldp * sizeof(int32_t), hipHostMallocDefault) ==
                 hipSuccess &&
             hipEventCreateWithFlags(&s->ring_ev[k], hipEventDisableTiming) == hipSuccess;
    const size_t nbal = (size_t)kRing * max_steps_per_tick * kBalWords;
    ok = ok && hipMalloc((void **)&s->bal_dev, nbal * sizeof(unsigned long long)) == hipSuccess &&
         hipHostMalloc((void **)&s->bal_host, nbal * sizeof(unsigned long long), hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc((void **)&s->err_host, kRing * sizeof(ErrWord), hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        serve_free(s);
        return hip_fail(ctx, hipErrorOutOfMemory);
    }
    *out = s;
    return CTSM_OK;
}

int ctsm_serve_submit(ctsm_server *s, int32_t ncol, int32_t npatch, int32_t npft, int32_t nsteps, int64_t *ticket)
{
    if (!s || !ticket || ncol < 1 || npatch < 0 || npft < 0 || nsteps < 1 || ncol > s->lay.ldc ||
        npatch > s->lay.ldp || npft > CTSM_MAX_PFT - s->pft_first)
        return CTSM_E_ARG;
    Member m;
    m.ncol = ncol;
    m.npatch = npatch;
    m.npft = npft;
    m.nsteps = m.left = nsteps;
    const int64_t t = s->next_ticket++;
    s->members[t] = m;
    s->queue.push_back(t);
    *ticket = t;
    return CTSM_OK;
}

int ctsm_serve_resume_member(ctsm_server *s, int32_t ncol, int32_t npatch, int32_t npft, int32_t left, int32_t col0,
                             int32_t pat0, int32_t pft0, int64_t *ticket)
{
    if (!s || !ticket || ncol < 1 || npatch < 0 || npft < 0 || left < 1 || col0 < 0 || pat0 < 0 || pft0 < 0)
        return CTSM_E_ARG;
    if (!s->cols.take(col0, ncol)) return CTSM_E_ARG;
    if (!s->pats.take(pat0, npatch)) {
        s->cols.release(col0, ncol);
        return CTSM_E_ARG;
    }
    if (!s->pfts.take(pft0 - s->pft_first, npft)) {
        s->cols.release(col0, ncol);
        s->pats.release(pat0, npatch);
        return CTSM_E_ARG;
    }
    Member m;
    m.ncol = ncol;
    m.npatch = npatch;
    m.npft = npft;
    m.nsteps = m.left = left;
    m.col0 = col0;
    m.pat0 = pat0;
    m.pft0 = pft0;
    m.state = CTSM_SERVE_RUNNING;
    const int64_t t = s->next_ticket++;
    s->members[t] = m;
    s->dirty = true;
    *ticket = t;
    return CTSM_OK;
}

int ctsm_serve_start(ctsm_server *s, int64_t ticket)
{
    if (!s) return CTSM_E_ARG;
    auto it = s->members.find(ticket);
    if (it == s->members.end() || it->second.state != CTSM_SERVE_ADMITTED) return CTSM_E_ARG;
    it->second.state = CTSM_SERVE_RUNNING;
    s->dirty = true;
    return CTSM_OK;
}

int ctsm_serve_poll(const ctsm_server *s, int64_t ticket, int32_t *state, int32_t *col0, int32_t *pat0, int32_t *pft0,
                    int32_t *left)
{
    if (!s || !state) return CTSM_E_ARG;
    auto it = s->members.find(ticket);
    if (it == s->members.end()) return CTSM_E_ARG;
    const Member &m = it->second;
    *state = (m.state == CTSM_SERVE_FINISHING) ? CTSM_SERVE_RUNNING : m.state;
    if (col0) *col0 = m.col0;
    if (pat0) *pat0 = m.pat0;
    if (pft0) *pft0 = m.pft0;
    if (left) *left = m.left;
    return CTSM_OK;
}

int ctsm_serve_release(ctsm_server *s, int64_t ticket)
{
    if (!s) return CTSM_E_ARG;
    auto it = s->members.find(ticket);
    if (it == s->members.end()) return CTSM_E_ARG;
    Member &m = it->second;
    if (m.state == CTSM_SERVE_RUNNING || m.state == CTSM_SERVE_FINISHING) return CTSM_E_ARG;
    if (m.state == CTSM_SERVE_QUEUED) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), ticket));
    } else {
        s->cols.release(m.col0, m.ncol);
        s->pats.release(m.pat0, m.npatch);
        s->pfts.release(m.pft0 - s->pft_first, m.npft);
    }
    s->members.erase(it);
    return CTSM_OK;
}

// Retires the oldest tick once its event has completed (or waits for it).
static int serve_retire(ctsm_server *s, bool wait)
{
    Tick &t = s->inflight.front();
    hipError_t he = wait ? hipEventSynchronize(t.ev) : hipEventQuery(t.ev);
    if (he == hipErrorNotReady) return 1;
    if (he != hipSuccess) return -hip_fail(s->ctx, he);
    const unsigned long long *w = s->bal_host + (size_t)t.slot * s->max_steps * kBalWords;
    auto worse = [](const unsigned long long *x, double &worst, int32_t &where) {
        double m;
        std::memcpy(&m, &x[0], sizeof(m));
        if (m > worst) {
            worst = m;
            where = (x[1] != 0x7fffffffull) ? (int32_t)x[1] : 0;
        }
    };
    for (int k = 0; k < t.steps; ++k) {
        const unsigned long long *x = w + (size_t)k * kBalWords;
        worse(x, s->worst_errsoi, s->worst_col);
        if (s->check_energy) {
            worse(x + 2, s->worst_errseb, s->worst_seb_patch);
            worse(x + 4, s->worst_errlon, s->worst_lon_patch);
        }
    }
    const ErrWord &e = s->err_host[t.slot];
    if (e.code != 0 && s->err_code == 0) {
        s->err_code = e.code;
        s->err_index = e.first_col;
    }
    for (auto &kv : s->members)
        if (kv.second.state == CTSM_SERVE_FINISHING && kv.second.finish_tick == t.id)
            kv.second.state = CTSM_SERVE_DONE;
    (void)hipEventDestroy(t.ev);
    s->inflight.pop_front();
    return 0;
}

int ctsm_serve_progress(ctsm_server *s, int32_t wait)
{
    if (!s || wait < 0 || wait > 2) return CTSM_E_ARG;
    ctsm_ctx *ctx = s->ctx;
    hipError_t he = hipSetDevice(ctx->device);
    if (he != hipSuccess) return hip_fail(ctx, he);
    const bool quiesce = wait == 2;
    // 1. retire completed ticks (wait 1: at least the oldest one; wait 2: all)
    bool waited = false;
    while (!s->inflight.empty()) {
        const int r = serve_retire(s, quiesce || (wait && !waited));
        if (r < 0) return -r;
        if (r > 0) break;
        waited = true;
    }
    if (s->err_code) {
        ctx->last_code = s->err_code;
        ctx->last_index = s->err_index;
        return s->err_code;
    }
    if (quiesce) return CTSM_OK;       // nothing admitted, nothing enqueued
    // 2. admit queued members first-come, while the pool has room
    while (!s->queue.empty()) {
        Member &m = s->members[s->queue.front()];
        const int64_t c0 = s->cols.alloc(m.ncol);
        if (c0 < 0) break;
        const int64_t p0 = s->pats.alloc(m.npatch);
        if (p0 < 0) {
            s->cols.release(c0, m.ncol);
            break;
        }
        const int64_t f0 = s->pfts.alloc(m.npft);
        if (f0 < 0) {
            s->cols.release(c0, m.ncol);
            s->pats.release(p0, m.npatch);
            break;
        }
        m.col0 = (int32_t)c0;
        m.pat0 = (int32_t)p0;
        m.pft0 = (int32_t)f0 + s->pft_first;
        m.state = CTSM_SERVE_ADMITTED;
        s->queue.pop_front();
    }
    // 3. one more tick over the running members, unless enough are in flight
    if ((int)s->inflight.size() >= kMaxInflight) return CTSM_OK;
    int32_t k = s->max_steps;
    bool any = false;
    for (auto &kv : s->members)
        if (kv.second.state == CTSM_SERVE_RUNNING) {
            k = std::min(k, kv.second.left);
            any = true;
        }
    if (!any) return CTSM_OK;
    if (s->dirty || s->cur < 0) {
        const int slot = (s->cur + 1) % kRing;
        // the slot's previous reader must be done before its buffers are refilled
        if (s->ring_used[slot] && (he = hipEventSynchronize(s->ring_ev[slot])) != hipSuccess)
            return hip_fail(ctx, he);
        int32_t nc = 0, np = 0;
        for (auto &kv : s->members) {
            const Member &m = kv.second;
            if (m.state != CTSM_SERVE_RUNNING) continue;
            for (int32_t i = 0; i < m.ncol; ++i) s->fc_host[slot][nc++] = s->lay.begc_alloc + m.col0 + i;
            for (int32_t i = 0; i < m.npatch; ++i) s->fp_host[slot][np++] = s->lay.begp_alloc + m.pat0 + i;
        }
        if (nc && (he = hipMemcpyAsync(s->fc_dev[slot], s->fc_host[slot], nc * sizeof(int32_t),
                                       hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
            return hip_fail(ctx, he);
        if (np && (he = hipMemcpyAsync(s->fp_dev[slot], s->fp_host[slot], np * sizeof(int32_t),
                                       hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
            return hip_fail(ctx, he);
        s->cur = slot;
        s->nc = nc;
        s->np = np;
        s->dirty = false;
    }
    const int slot = s->cur;
    unsigned long long *bal = s->bal_dev + (size_t)slot * s->max_steps * kBalWords;
    for (int32_t st = 0; st < k; ++st) {
        // each member's bounds are its own patches: members finished in
        // an earlier tick and not yet copied out keep their results
        int rc = land_step_full(ctx, &s->lay, &s->col, &s->pch, &s->can, &s->sf, s->have_sw ? &s->sw : nullptr,
                                (s->have_sw || s->have_dr) ? &s->hprm : nullptr, s->have_wt ? &s->wt : nullptr,
                                s->have_dr ? &s->dr : nullptr, s->nc, s->fc_dev[slot], s->np, s->fp_dev[slot],
                                s->dtime, true);
        if (rc) return rc;
        // EnergyBalanceCheck after every step (BalanceCheckMod.F90:966-1003, 1101-1119)
        unsigned long long *b = bal + (size_t)st * kBalWords;
        launch_balance(ctx->stream, s->lay, ctx->cfg, s->sf.errsoi_col, s->nc, s->fc_dev[slot], b);
        if (s->check_energy)
            launch_energy_balance(ctx->stream, s->lay, ctx->cfg, s->col, s->pch, s->can, s->sf, s->np,
                                  s->fp_dev[slot], nullptr, nullptr, b + 2, ctx->err);
    }
    // the tick's results and error word come back with it; the error word is
    // then cleared for the next tick
    if ((he = hipMemcpyAsync(s->bal_host + (size_t)slot * s->max_steps * kBalWords, bal,
                             (size_t)kBalWords * k * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, ctx->stream)) != hipSuccess ||
        (he = hipMemcpyAsync(&s->err_host[slot], ctx->err, sizeof(ErrWord), hipMemcpyDeviceToHost, ctx->stream)) !=
            hipSuccess)
        return hip_fail(ctx, he);
    hipLaunchKernelGGL(k_reset_err, dim3(1), dim3(1), 0, ctx->stream, ctx->err);
    Tick t;
    t.id = s->next_tick++;
    t.slot = slot;
    t.steps = k;
    if ((he = hipEventCreateWithFlags(&t.ev, hipEventDisableTiming)) != hipSuccess ||
        (he = hipEventRecord(t.ev, ctx->stream)) != hipSuccess ||
        (he = hipEventRecord(s->ring_ev[slot], ctx->stream)) != hipSuccess)
        return hip_fail(ctx, he);
    s->ring_used[slot] = true;
    s->inflight.push_back(t);
    s->column_steps += (int64_t)k * s->nc;
    for (auto &kv : s->members) {
        Member &m = kv.second;
        if (m.state != CTSM_SERVE_RUNNING) continue;
        m.left -= k;
        if (m.left == 0) {
            m.state = CTSM_SERVE_FINISHING;
            m.finish_tick = t.id;
            s->dirty = true;
        }
    }
    return finish_launch(ctx);
}

int ctsm_serve_stats(const ctsm_server *s, int64_t *column_steps, double *worst_errsoi, int32_t *worst_col,
                     double *worst_errseb, int32_t *seb_patch, double *worst_errlon, int32_t *lon_patch,
                     int32_t *inflight)
{
    if (!s) return CTSM_E_ARG;
    if (worst_errseb) *worst_errseb = s->worst_errseb;
    if (seb_patch) *seb_patch = s->worst_seb_patch;
    if (worst_errlon) *worst_errlon = s->worst_errlon;
    if (lon_patch) *lon_patch = s->worst_lon_patch;
    if (column_steps) *column_steps = s->column_steps;
    if (worst_errsoi) *worst_errsoi = s->worst_errsoi;
    if (worst_col) *worst_col = s->worst_col;
    if (inflight) *inflight = (int32_t)s->inflight.size();
    return CTSM_OK;
}

int ctsm_serve_destroy(ctsm_server *s)
{
    if (!s) return CTSM_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    serve_free(s);
    return CTSM_OK;
}
