// soilwater.hip -- moisture-form Richards solve on gfx950.
//
// Replaces soilwater_moisture_form (SoilWaterMovementMod.F90:976-1440).  One
// column per lane: the column's <= 20 layers live in VGPRs/AGPRs for the
// whole adaptive sub-step loop, so a trial touches no memory except the
// column's inputs loaded once.  Each trial is two sweeps over the layers:
//
//   forward : Clapp-Hornberger hydraulic properties (:1444-1556, CH78
//             soil_hk / soil_suction), Darcy fluxes and their derivatives
//             (:1560-1866), the RHS / LHS rows (:1871-2012), and the
//             elimination step of LAPACK dgtsv for the previous row, all
//             streamed one layer behind the property lookahead;
//   backward: dgtsv back substitution fused with the error estimate
//             (:1307-1351);
//
// then the accept / halve / double decision of :1353-1397.  Every value is
// computed by the reference's expression in the reference's order, so the
// fusion changes no rounding.  Lanes of a wavefront retire independently;
// the column order in the filter (sorted by layer count by the caller) is
// what keeps a wavefront's lanes doing the same amount of work.
#include "common.h"
#include "bucket_order.h"
#include "trace.h"
#include "soilwater_dev.h"
#include "richards_group.h"

namespace ctsm {

template <bool WT>
__global__ void __launch_bounds__(64)
k_soilwater(View V, Params P, ctsm_col_state S, int n, const int32_t *__restrict__ filter,
            double dtime, ErrWord *err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (!col_in_range(V, c)) { report_error(err, c, CTSM_E_ARG); return; }
    const int rc = richards_column<MAX_NLEVSOI, WT>(V, P, S, c, dtime);
    if (rc) report_error(err, c, richards_error_code(rc));
}

// SetSoilWaterFractions (SoilHydrologyMod.F90:240-252), excess_ice = 0.
__global__ void __launch_bounds__(256)
k_soilwater_fractions(View V, ctsm_col_state S, int n, const int32_t *__restrict__ filter)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (!col_in_range(V, c)) return;
    for (int j = 1; j <= V.nlevsoi; ++j) {
        const double ws = S.watsat[V.soi(c, j)];
        const double dz_ext = S.dz[V.sno(c, j)] + 0.0 / DENICE;
        const double vol_ice = fmin90(ws, (S.h2osoi_ice[V.sno(c, j)] + 0.0) / (dz_ext * DENICE));
        S.eff_porosity[V.soi(c, j)] = fmax90(0.01, ws - vol_ice);
        S.icefrac[V.soi(c, j)] = fmin90(1.0, vol_ice / ws);
    }
}

// G lanes per column (richards_group.h).  Groups are aligned to G lanes, so a
// column's lanes retire together.  The solve is latency-bound (a trial is a
// serial elimination chain), so for G = 8 the register budget is capped at 168
// VGPRs for three waves per SIMD: the few values that then live in scratch
// (10 VGPRs) cost less than the latency the extra wave hides.  Measured on the
// columns ordered by their previous sub-step count, 1,048,576 uneven columns:
// 5.07 ms at four waves (77 VGPRs spilled), 4.37 at three, 5.17 at two.
template <int G> struct RichardsWaves { static constexpr int value = 1; };
template <> struct RichardsWaves<8> { static constexpr int value = 3; };

template <int G, bool WT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RichardsWaves<G>::value)))
k_soilwater_g(View V, Params P, ctsm_col_state S, int n, const int32_t *__restrict__ filter, double dtime,
              ErrWord *err)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / G, g = t % G;
    if (i >= n) return;
    const int c = filter ? filter[i] : V.begc + i;
    if (!col_in_range(V, c)) { if (g == 0) report_error(err, c, CTSM_E_ARG); return; }
    const int rc = richards_group<G, WT>(V, P, S, c, dtime, g);
    if (rc && g == 0) report_error(err, c, richards_error_code(rc));
}

// Group size chosen per wavefront (richards_lanes = 0).  A one-wave block
// takes 32 consecutive columns of the filter and the deepest bedrock index
// among them; groups of 3 layers per lane then need 2 lanes per column up to
// 6 layers, 4 up to 12 and 8 up to 20, and the block runs its columns in
// passes of 64 / lanes.  A batch stored by depth gives most blocks the
// smallest group its columns allow, so no lane holds only empty layers; the
// results do not depend on the group size.  Built for three waves per SIMD
// (168 VGPRs): at four (128 VGPRs, 120 spilled) the solve of 1,048,576
// uneven columns took 8.3 ms, at three 7.4, at two 8.1; with the columns
// ordered by sub-step count and depth class (SubstepDepthKey) 6.3-7.0 ms.
// A deep block runs its 32 columns in four passes of 8, so the 8-lane kernel
// (4.4 ms with its order) stays the default.
constexpr int ADAPT_SPAN = 32;
constexpr int ADAPT_K = 3;
template <bool WT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
k_soilwater_adaptive(View V, Params P, ctsm_col_state S, int n, const int32_t *__restrict__ filter, double dtime,
                     ErrWord *err)
{
    const int lane = threadIdx.x;
    const int i0 = blockIdx.x * ADAPT_SPAN;
    const int i1 = min(i0 + ADAPT_SPAN, n);
    int nl = 0;
    if (lane < ADAPT_SPAN && i0 + lane < i1) {
        const int c = filter ? filter[i0 + lane] : V.begc + i0 + lane;
        if (col_in_range(V, c)) nl = S.nbedrock[V.c1(c)];
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) nl = max(nl, __shfl_xor(nl, m));
    const int gsz = (nl <= 2 * ADAPT_K) ? 2 : (nl <= 4 * ADAPT_K) ? 4 : 8;
    const int per = 64 / gsz;
    for (int base = i0; base < i1; base += per) {
        const int i = base + lane / gsz, g = lane % gsz;
        if (i < i1) {
            const int c = filter ? filter[i] : V.begc + i;
            if (!col_in_range(V, c)) {
                if (g == 0) report_error(err, c, CTSM_E_ARG);
            } else {
                const int rc = richards_group<0, WT, ADAPT_K>(V, P, S, c, dtime, g, gsz);
                if (rc && g == 0) report_error(err, c, richards_error_code(rc));
            }
        }
    }
}

}  // namespace ctsm

namespace ctsm {
template <bool WT>
void launch_soilwater_k(hipStream_t st, const View &V, const Params &P, const ctsm_col_state &S, int n,
                        const int32_t *filter, double dtime, ErrWord *err, int lanes, int block);

void launch_soilwater(hipStream_t st, const View &V, const Params &P, const ctsm_col_state &S, int n,
                      const int32_t *filter, double dtime, ErrWord *err, int lanes, int block)
{
    if (n <= 0) return;
    const GptlRange range("hydro_without_drainage");
    if (P.lower_bc == CTSM_BC_WATERTABLE) launch_soilwater_k<true>(st, V, P, S, n, filter, dtime, err, lanes, block);
    else launch_soilwater_k<false>(st, V, P, S, n, filter, dtime, err, lanes, block);
}

template <bool WT>
void launch_soilwater_k(hipStream_t st, const View &V, const Params &P, const ctsm_col_state &S, int n,
                        const int32_t *filter, double dtime, ErrWord *err, int lanes, int block)
{
    if (lanes == 0) {
        hipLaunchKernelGGL(k_soilwater_adaptive<WT>, dim3((unsigned)((n + ADAPT_SPAN - 1) / ADAPT_SPAN)), dim3(64), 0,
                           st, V, P, S, n, filter, dtime, err);
        return;
    }
    if (block <= 0 || block > 256) block = (lanes == 1) ? 64 : 128;
    if (lanes != 2 && lanes != 4 && lanes != 8) block = 64;     /* k_soilwater: __launch_bounds__(64) */
    const long threads = (long)n * (lanes > 0 ? lanes : 1);
    const dim3 grid((unsigned)((threads + block - 1) / block)), blk(block);
    switch (lanes) {
    case 2: hipLaunchKernelGGL((k_soilwater_g<2, WT>), grid, blk, 0, st, V, P, S, n, filter, dtime, err); break;
    case 4: hipLaunchKernelGGL((k_soilwater_g<4, WT>), grid, blk, 0, st, V, P, S, n, filter, dtime, err); break;
    case 8: hipLaunchKernelGGL((k_soilwater_g<8, WT>), grid, blk, 0, st, V, P, S, n, filter, dtime, err); break;
    default: hipLaunchKernelGGL(k_soilwater<WT>, grid, blk, 0, st, V, P, S, n, filter, dtime, err); break;
    }
}
// ---- the Richards solve's column order ----------------------------------
// The solve sub-steps adaptively, and the lanes of a wave iterate until the
// last of its columns is done: columns that need few sub-steps wait for a
// neighbour that needs many.  The previous step's count predicts the next,
// so a step can solve its columns sorted by it (bucket_order.h, BO_KEYS
// buckets, the last one open-ended) and pass the sorted column list to the
// solve as its filter.  Each column's result does not depend on the order its
// wave was formed in.
struct SubstepKey {
    View V;
    ctsm_col_state S;
    const int32_t *filter;
    __device__ int32_t value(int i) const { return filter ? filter[i] : V.begc + i; }
    __device__ int key(int i) const
    {
        const int c = value(i);
        if (!col_in_range(V, c)) return 0;
        const double ns = S.num_substeps[V.c1(c)];
        if (!(ns > 0.0)) return 0;
        return ns >= (double)(BO_KEYS - 1) ? BO_KEYS - 1 : (int)ns;
    }
};

// For the adaptive-group kernel: the sub-step count (up to 20) and, within
// it, the group size the column's depth needs (2, 4 or 8 lanes), so the 32
// consecutive columns a wave takes share both.
struct SubstepDepthKey {
    SubstepKey k;
    __device__ int32_t value(int i) const { return k.value(i); }
    __device__ int key(int i) const
    {
        const int c = value(i);
        if (!col_in_range(k.V, c)) return 0;
        const int ns = k.key(i) < 20 ? k.key(i) : 20;
        const int nl = k.S.nbedrock[k.V.c1(c)];
        return ns * 3 + (nl <= 2 * ADAPT_K ? 0 : nl <= 4 * ADAPT_K ? 1 : 2);
    }
};

size_t richards_order_bytes(int n) { return (size_t)n * sizeof(int32_t) + bucket_order_table_bytes(n); }

// The filter entries sorted by num_substeps into ws (richards_order_bytes(n));
// returns the sorted column list (inside ws).  lanes: the Richards kernel's
// lanes per column (0: the adaptive-group kernel, which also sorts by depth).
const int32_t *launch_richards_order(hipStream_t st, const View &V, const ctsm_col_state &S, int n,
                                     const int32_t *filter, void *ws, int lanes)
{
    int32_t *order = (int32_t *)ws;
    if (lanes == 0)
        launch_bucket_order(st, SubstepDepthKey{SubstepKey{V, S, filter}}, n, nullptr, order + n, order);
    else
        launch_bucket_order(st, SubstepKey{V, S, filter}, n, nullptr, order + n, order);
    return order;
}

void launch_soilwater_fractions(hipStream_t st, const View &V, const ctsm_col_state &S, int n,
                                const int32_t *filter)
{
    if (n <= 0) return;
    const int bs = 256;
    hipLaunchKernelGGL(k_soilwater_fractions, dim3((n + bs - 1) / bs), dim3(bs), 0, st, V, S, n, filter);
}
}  // namespace ctsm
