// history.hip -- history buffers of device-resident state on gfx950.
//
// hist_update_hbuf (histFileMod.F90:1362-2169, the branch that keeps a field at
// the resolution of its source), hfields_normalize (:2302-2358) and
// hfields_zero (:2361-2381).  Every element of every buffer is independent, so
// a launch covers up to kHistBatch fields (the descriptors travel by value in
// the kernel arguments, which keeps the calls capturable).  Each block takes
// 256 consecutive points of one level of one field: the block finds its
// (field, level, chunk) once, by a binary search over the fields' block
// offsets, and its lanes read the source row with unit stride -- no per-element
// index division.  HBM-bound: 8 B read + 12 B read and written per element.
#include "common.h"

namespace ctsm {
namespace hist {

constexpr int kHistBatch = 24;
constexpr double SPVAL = 1.e36;

constexpr int kRow = 256;   // points per block

struct Batch {
    ctsm_hist_field f[kHistBatch];
    int boff[kHistBatch + 1];   // first block of each field
    int chunks[kHistBatch];     // blocks per level row
    int n;
};

// The block's field, and its element range: point k0.. of level j.
struct Slot {
    int fi;
    long long row;   // j * n1d
    int k0, j;
};

__device__ __forceinline__ Slot find_slot(const Batch &B, int b)
{
    int lo = 0, hi = B.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.boff[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const int r = b - B.boff[lo];
    const int j = r / B.chunks[lo];
    Slot s;
    s.fi = lo;
    s.j = j;
    s.k0 = (r - j * B.chunks[lo]) * kRow;
    s.row = (long long)j * B.f[lo].n1d;
    return s;
}

__global__ void __launch_bounds__(256) k_hist_update(Batch B)
{
    const Slot s = find_slot(B, blockIdx.x);
    const ctsm_hist_field &F = B.f[s.fi];
    const int k = s.k0 + threadIdx.x;
    if (k < F.n1d) {
        const long long i = s.row + k;
        const bool valid = !F.active || F.active[k] != 0;
        const double x = F.field[k + s.j * F.ld];
        double h = F.hbuf[i];
        int na = F.nacs[i];
        switch (F.avgflag) {
        case 'I':
            h = valid ? ((x != SPVAL) ? x : SPVAL) : SPVAL;
            na = 1;
            break;
        case 'A':
        case 'S':
            if (valid && x != SPVAL) {
                if (na == 0) h = 0.0;
                h = h + x;
                na = na + 1;
            } else if (na == 0) {
                h = SPVAL;
            }
            break;
        case 'X':
            if (valid && x != SPVAL) {
                if (na == 0) h = -1.e50;
                h = fmax90(h, x);
            } else if (na == 0) {
                h = SPVAL;
            }
            na = 1;
            break;
        default:  // 'M'
            if (valid && x != SPVAL) {
                if (na == 0) h = +1.e50;
                h = fmin90(h, x);
            } else if (na == 0) {
                h = SPVAL;
            }
            na = 1;
            break;
        }
        F.hbuf[i] = h;
        F.nacs[i] = na;
    }
}

// hfields_normalize: only the 'A' buffers are divided, by float(nacs)
__global__ void __launch_bounds__(256) k_hist_normalize(Batch B)
{
    const Slot s = find_slot(B, blockIdx.x);
    const ctsm_hist_field &F = B.f[s.fi];
    const int k = s.k0 + threadIdx.x;
    if (k >= F.n1d || F.avgflag != 'A') return;
    const long long i = s.row + k;
    const int na = F.nacs[i];
    if (na != 0) F.hbuf[i] = F.hbuf[i] / (double)(float)na;
}

__global__ void __launch_bounds__(256) k_hist_zero(Batch B)
{
    const Slot s = find_slot(B, blockIdx.x);
    const ctsm_hist_field &F = B.f[s.fi];
    const int k = s.k0 + threadIdx.x;
    if (k >= F.n1d) return;
    const long long i = s.row + k;
    F.hbuf[i] = 0.0;
    F.nacs[i] = 0;
}

}  // namespace hist

// op: 0 update, 1 normalize, 2 zero.  The descriptors were validated by the caller.
void launch_hist(hipStream_t st, const ctsm_hist_field *fields, int nfields, int op)
{
    using namespace hist;
    for (int b = 0; b < nfields; b += kHistBatch) {
        Batch B{};
        B.n = 0;
        B.boff[0] = 0;
        for (int f = b; f < std::min(nfields, b + kHistBatch); ++f) {
            if ((long long)fields[f].n1d * fields[f].num2d == 0) continue;   // nothing to do
            const int chunks = (fields[f].n1d + kRow - 1) / kRow;
            B.f[B.n] = fields[f];
            B.chunks[B.n] = chunks;
            B.boff[B.n + 1] = B.boff[B.n] + chunks * fields[f].num2d;
            ++B.n;
        }
        const int blocks = B.boff[B.n];
        if (blocks <= 0) continue;
        if (op == 0) hipLaunchKernelGGL(k_hist_update, dim3(blocks), dim3(kRow), 0, st, B);
        else if (op == 1) hipLaunchKernelGGL(k_hist_normalize, dim3(blocks), dim3(kRow), 0, st, B);
        else hipLaunchKernelGGL(k_hist_zero, dim3(blocks), dim3(kRow), 0, st, B);
    }
}

}  // namespace ctsm
