// phs.hip -- the soil side of plant hydraulic stress (use_hydrstress) on gfx950.
//
// Two per-element loops of the PHS path:
//   * the root-soil interface conductance at the head of
//     PhotosynthesisHydraulicStress (PhotosynthesisMod.F90:3062-3114), one lane
//     per patch, the soil layers in order inside the lane, with the root
//     segment's Weibull vulnerability curve plc (:5167-5195);
//   * Compute_EffecRootFrac_And_VertTranSink_HydStress
//     (SoilWaterPlantSinkMod.F90:236-328), one lane per column: the patch
//     fluxes k_soil_root (smp - vegwp(root) - z) summed per layer over the
//     column's patches in ascending order (the reference's j-outer, p-inner
//     loop gives each layer that order), and each patch's negative part over
//     the layers in ascending order.
// Both are HBM-bound streams: every lane reads and writes its own elements once;
// patch-major layouts make consecutive lanes touch consecutive addresses.
#include "common.h"

namespace ctsm {
namespace phs {

constexpr double RPI = 3.14159265358979323846;   // SHR_CONST_PI
constexpr double C_TO_B = 2.0;                     // g biomass / g C (PhotosynthesisMod.F90:2911)
constexpr double CROOT_LATERAL_LENGTH = 0.25;      // m (:2910)

// rows of the device parameter table, CTSM_MAX_PFT values each
enum { T_KRMAX = 0, T_RADIUS = 1, T_DENSITY = 2, T_FROOT_LEAF = 3, T_PSI50 = 4, T_CK = 8, T_ROWS = 12 };

struct PView {
    int begc, ldc, begp, ldp, nlevsno, nlevsoi;
    __device__ __forceinline__ long c1(int c) const { return (long)(c - begc); }
    __device__ __forceinline__ long p1(int p) const { return (long)(p - begp); }
    __device__ __forceinline__ long soi(int c, int j) const { return (long)(c - begc) + (long)(j - 1) * ldc; }
    __device__ __forceinline__ long sno(int c, int j) const { return (long)(c - begc) + (long)(j + nlevsno - 1) * ldc; }
    __device__ __forceinline__ long soip(int p, int j) const { return (long)(p - begp) + (long)(j - 1) * ldp; }
    __device__ __forceinline__ long veg(int p, int k) const { return (long)(p - begp) + (long)(k - 1) * ldp; }
};

// plc, vegetation_weibull branch: 2**(-(x/psi50)**ck), cut to zero below 0.005
__device__ __forceinline__ double plc(double x, double psi50, double ck)
{
    double v = cm_pow(2.0, -cm_pow(x / psi50, ck));
    if (v < 0.005) v = 0.0;
    return v;
}

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

}  // namespace phs

void launch_phs_conductance(hipStream_t st, const ctsm_layout &L, const ctsm_config &cfg, const ctsm_phs_state &S,
                            const double *table, int npft, int n, const int32_t *filter, ErrWord *err)
{
    if (n <= 0) return;
    const phs::PView V{L.begc_alloc, L.ldc, L.begp_alloc, L.ldp, cfg.nlevsno, cfg.nlevsoi};
    hipLaunchKernelGGL(phs::k_phs_conductance, dim3((n + 255) / 256), dim3(256), 0, st, V, S, table, npft, n, filter,
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
