// phs.hip -- the soil side of plant hydraulic stress (use_hydrstress) on gfx950.
//
// Per-element loops of the PHS path:
//   * the root-soil interface conductance at the head of
//     PhotosynthesisHydraulicStress (PhotosynthesisMod.F90:3062-3114), one lane
//     per patch, the soil layers in order inside the lane, with the root
//     segment's Weibull vulnerability curve plc (:5167-5195);
//   * Compute_EffecRootFrac_And_VertTranSink_HydStress
//     (SoilWaterPlantSinkMod.F90:236-328), one lane per column: the patch
//     fluxes k_soil_root (smp - vegwp(root) - z) summed per layer over the
//     column's patches in ascending order (the reference's j-outer, p-inner
//     loop gives each layer that order), and each patch's negative part over
//     the layers in ascending order;
//   * calcstress (PhotosynthesisMod.F90:4490-4710), one lane per patch: the
//     vegetation water potential's four-segment Newton iteration.
// The first two are HBM-bound streams: every lane reads and writes its own elements once;
// patch-major layouts make consecutive lanes touch consecutive addresses.
#include "common.h"
#include "phs_dev.h"

namespace ctsm {
namespace phs {

struct PView {
    int begc, ldc, begp, ldp, nlevsno, nlevsoi;
    __device__ __forceinline__ long c1(int c) const { return (long)(c - begc); }
    __device__ __forceinline__ long p1(int p) const { return (long)(p - begp); }
    __device__ __forceinline__ long soi(int c, int j) const { return (long)(c - begc) + (long)(j - 1) * ldc; }
    __device__ __forceinline__ long sno(int c, int j) const { return (long)(c - begc) + (long)(j + nlevsno - 1) * ldc; }
    __device__ __forceinline__ long soip(int p, int j) const { return (long)(p - begp) + (long)(j - 1) * ldp; }
    __device__ __forceinline__ long veg(int p, int k) const { return (long)(p - begp) + (long)(k - 1) * ldp; }
};

__global__ void __launch_bounds__(256) k_phs_conductance(PView V, ctsm_phs_state S, const double *__restrict__ T,
                                                         int npft, int n, const int32_t *__restrict__ filter,
                                                         ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = filter ? filter[i] : V.begp + i;
    if (p < V.begp || p >= V.begp + V.ldp) { report_error(err, p, CTSM_E_ARG); return; }
    const long pp = V.p1(p);
    const int c = S.column[pp], ivt = S.itype[pp];
    if (c < V.begc || c >= V.begc + V.ldc || ivt < 0 || ivt >= npft) { report_error(err, p, CTSM_E_ARG); return; }
    const double froot = S.froot_carbon[pp];
    const double lai = S.tsai[pp] + S.tlai[pp];
    const double radius = T[T_RADIUS * CTSM_MAX_PFT + ivt];
    const double density = T[T_DENSITY * CTSM_MAX_PFT + ivt];
    const double froot_leaf = T[T_FROOT_LEAF * CTSM_MAX_PFT + ivt];
    const double krmax = T[T_KRMAX * CTSM_MAX_PFT + ivt];
    const double psi50 = T[(T_PSI50 + 3) * CTSM_MAX_PFT + ivt];   // root segment
    const double ck = T[(T_CK + 3) * CTSM_MAX_PFT + ivt];
    for (int j = 1; j <= V.nlevsoi; ++j) {
        const double rf = S.rootfr[V.soip(p, j)];
        const double zj = S.z[V.sno(c, j)];
        double rbd = C_TO_B * froot * rf / S.dz[V.sno(c, j)];
        rbd = fmax90(C_TO_B * 1.0, rbd);
        const double cross = RPI * (radius * radius);
        const double rld = rbd / (density * cross);
        const double rai = lai * froot_leaf * rf;
        const double r_soil = sqrt(1.0 / (RPI * rld));
        double soil_c = fmin90(S.hksat[V.soi(c, j)], S.hk_l[V.soi(c, j)]) / (1.e3 * r_soil);
        const double fs = plc(S.smp_l[V.soi(c, j)], psi50, ck);
        double root_c = (fs * rai * krmax) / (CROOT_LATERAL_LENGTH + zj);
        soil_c = fmax90(soil_c, 1.e-16);
        root_c = fmax90(root_c, 1.e-16);
        S.root_conductance[V.soip(p, j)] = root_c;
        S.soil_conductance[V.soip(p, j)] = soil_c;
        const double rs_resis = 1.0 / soil_c + 1.0 / root_c;
        S.k_soil_root[V.soip(p, j)] = (rai * rf > 0.0 && j > 1) ? 1.0 / rs_resis : 0.0;
    }
}

__global__ void __launch_bounds__(256) k_phs_root_sink(PView V, ctsm_phs_state S, int n,
                                                       const int32_t *__restrict__ filter, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (c < V.begc || c >= V.begc + V.ldc) { report_error(err, c, CTSM_E_ARG); return; }
    const long cc = V.c1(c);
    const int pi = S.patchi[cc], pf = S.patchf[cc];
    if (pf >= pi && (pi < V.begp || pf >= V.begp + V.ldp)) { report_error(err, c, CTSM_E_ARG); return; }
    double r[MAX_NLEVSOI];
#pragma unroll
    for (int j = 0; j < MAX_NLEVSOI; ++j) r[j] = 0.0;
    for (int p = pi; p <= pf; ++p) {
        const long q = V.p1(p);
        double redist = 0.0;
        const double w = S.wtcol[q];
        if (S.active[q] != 0 && S.frac_veg_nosno[q] > 0 && w > 0.0) {
            const double psi_root = S.vegwp[V.veg(p, 4)];
#pragma unroll
            for (int j = 1; j <= MAX_NLEVSOI; ++j) {
                if (j > V.nlevsoi) continue;
                const double grav2 = S.z[V.sno(c, j)] * 1000.0;
                const double flux = S.k_soil_root[V.soip(p, j)] * (S.smp_l[V.soi(c, j)] - psi_root - grav2);
                if (flux < 0.0) redist = redist + flux;
                r[j - 1] = r[j - 1] + flux * w;
            }
        }
        S.qflx_hydr_redist[q] = redist;
    }
    double neg = 0.0;
#pragma unroll
    for (int j = 1; j <= MAX_NLEVSOI; ++j) {
        if (j > V.nlevsoi) continue;
        S.qflx_rootsoi[V.soi(c, j)] = r[j - 1];
        if (r[j - 1] < 0.0) neg = neg + r[j - 1];
    }
    S.qflx_phs_neg[cc] = neg;
}

// Three waves per SIMD (168 VGPRs, 124 B of scratch per lane): 4.52 ms against 4.93 ms
// at the compiler's own choice of two, for 655,821 patches.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) k_phs_calcstress(PView V, ctsm_phs_stress_state S, const double *__restrict__ T,
                                                       int npft, int n, const int32_t *__restrict__ filter,
                                                       ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int p = filter ? filter[i] : V.begp + i;
    if (p < V.begp || p >= V.begp + V.ldp) { report_error(err, p, CTSM_E_ARG); return; }
    const long pp = V.p1(p);
    const int c = S.column[pp], ivt = S.itype[pp];
    if (c < V.begc || c >= V.begc + V.ldc || ivt < 0 || ivt >= npft) { report_error(err, p, CTSM_E_ARG); return; }
    Leaf P;
    P.laisun = S.laisun[pp]; P.laisha = S.laisha[pp]; P.elai = S.elai[pp]; P.esai = S.esai[pp];
    P.tsai = S.tsai[pp]; P.htop = S.htop[pp]; P.fdry = S.fdry[pp]; P.thm = S.thm[pp];
    P.rho = S.forc_rho[V.c1(c)]; P.pbot = S.forc_pbot[V.c1(c)]; P.gb = S.gb_mol[pp];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        P.kmax[k] = T[(T_KMAX + k) * CTSM_MAX_PFT + ivt];
        P.psi50[k] = T[(T_PSI50 + k) * CTSM_MAX_PFT + ivt];
        P.ck[k] = T[(T_CK + k) * CTSM_MAX_PFT + ivt];
    }
    P.k = S.k_soil_root + V.soip(p, 1);
    P.smp = S.smp_l + V.soi(c, 1);
    P.z = S.z + V.sno(c, 1);
    P.sk = V.ldp;
    P.ss = V.ldc;
    P.n = V.nlevsoi;
    const double qsatl = S.qsatl[pp], qaf = S.qaf[pp];
    const double gs_sun_in = S.gs_mol_sun[pp], gs_sha_in = S.gs_mol_sha[pp];
    double x[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = S.vegwp[V.veg(p, k + 1)];

    double bsun, bsha, tran = 0.0;
    const bool night = calcstress(P, x, gs_sun_in, gs_sha_in, qsatl, qaf, bsun, bsha, tran);
    if (night) S.qflx_tran_veg[pp] = tran;
    const bool pd = night && S.local_morning[pp] != 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        S.vegwp[V.veg(p, k + 1)] = x[k];
        S.vegwp_pd[V.veg(p, k + 1)] = pd ? x[k] : SPVAL;
    }
    S.bsun[pp] = bsun;
    S.bsha[pp] = bsha;
}

}  // namespace phs

void launch_phs_conductance(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_phs_state &S,
                            const double *table, int npft, int n, const int32_t *filter, ErrWord *err)
{
    if (n <= 0) return;
    const phs::PView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, cfg.nlevsno, cfg.nlevsoi};
    hipLaunchKernelGGL(phs::k_phs_conductance, dim3((n + 255) / 256), dim3(256), 0, st, V, S, table, npft, n, filter,
                       err);
}

void launch_phs_calcstress(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg,
                           const ctsm_phs_stress_state &S, const double *table, int npft, int n,
                           const int32_t *filter, ErrWord *err)
{
    if (n <= 0) return;
    const phs::PView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, cfg.nlevsno, cfg.nlevsoi};
    hipLaunchKernelGGL(phs::k_phs_calcstress, dim3((n + 63) / 64), dim3(64), 0, st, V, S, table, npft, n, filter,
                       err);
}

void launch_phs_root_sink(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_phs_state &S,
                          int n, const int32_t *filter, ErrWord *err)
{
    if (n <= 0) return;
    const phs::PView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, cfg.nlevsno, cfg.nlevsoi};
    hipLaunchKernelGGL(phs::k_phs_root_sink, dim3((n + 255) / 256), dim3(256), 0, st, V, S, n, filter, err);
}

}  // namespace ctsm
