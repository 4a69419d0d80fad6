// Half-row batched LU for the weight systems of 97 .. 112 unknowns (the register-row classes of
// rowlu.h stop at 96: a row of 112 + 4 doubles does not fit a thread's registers).  One
// work-group of 4 wavefronts per system: thread t holds half h = t / 128 of row r = t % 128 of
// [A | B] in registers, columns 56 h .. 56 h + 55 (half 0: A's columns 0..55; half 1: A's
// columns 56..111, then the four right-hand sides: 56 and 60 registers pairs, which spill less
// than a split at 64 and took config C's half-row systems from 73.6 to 69 ms).  The elimination follows rowlu.h (partial
// pivoting on the largest |a_ik|, the lower row on ties; rows are not swapped, each keeps the step
// at which it became the pivot), with the columns of a block of 4 steps in one known half:
//   - the pivot: the two wavefronts of the half holding column k reduce |a_ik| (DPP steps and a
//     ballot) and post their candidates with the signed a_ik; one barrier;
//   - every thread picks the winner; the pivot row's halves post its columns from k on and the
//     half holding column k posts each unused row's multiplier; one barrier;
//   - the unused rows update their half's columns past k with register FMAs.
// The back substitution runs in half 1 for the columns 56.. (where the right-hand sides are), then
// hands the partial sums to half 0 through LDS; the refinement's forward substitution the other
// way round.  Its residual b - A x re-evaluates A from the coordinates, each half its own
// columns.  Same systems, pivot rule and outputs as the register-row and LDS kernels.
#include <hip/hip_runtime.h>

#include "rowlu.h"

namespace serbf {
namespace {

constexpr int kHalf = 56;                        // columns of A per half row (a multiple of 8)
constexpr int kHMaxA = 112;                      // unknowns of the class
constexpr int kHA1 = kHMaxA - kHalf;             // A's columns in half 1 (56)
constexpr int kHC1 = kHA1 + kRhs;                // half 1's columns with the right-hand sides (60)
constexpr int kHN = kHalf > kHC1 ? kHalf : kHC1; // register row length (60)
constexpr int kHBlocks = kHMaxA / kRowChunk;     // blocks of 4 elimination steps (28)
constexpr int kHSwitch = kHalf / kRowChunk;      // first block of half 1 (14)
// the assembly loads 8-column blocks of A up to kHN: both halves' A columns must fit them
static_assert(kHalf % 8 == 0 && kHA1 <= 8 * (kHN / 8), "half-row split");

struct HalfSys {
    double *pr, *lm, *dinv, *red, *redv, *xfer, *rb;
    int *plane, *redi;
    int r, h, lane, wv, m;
};

// A columns held by half H (compile time)
template <int H>
constexpr int half_len() { return H ? kHC1 : kHalf; }

// steps 4 CH .. 4 CH + 3 of the elimination; false once k reaches m or a pivot is zero
template <int CH>
__device__ __forceinline__ bool hr_factor(double (&row)[kHN], int &slot, const HalfSys &S, int &info)
{
    constexpr int B0 = kRowChunk * CH;
    constexpr int HK = B0 >= kHalf ? 1 : 0;   // the half holding columns B0 .. B0 + 3
    constexpr int LB0 = B0 - HK * kHalf;      // their local index there
    for (int kk = 0; kk < kRowChunk; kk++) {
        const int k = B0 + kk;
        if (k >= S.m) return false;
        if (S.h == HK) {
            const double v = sel4<LB0>(row, kk);
            const double mine = slot < 0 ? fabs(v) : -1.0;
            const double best = wave_max(mine);
            const unsigned long long hit = __ballot(slot < 0 && mine == best);
            const bool nan = __ballot(slot < 0 && mine != mine) != 0;
            const int q = S.wv - 2 * HK;   // 0, 1: rows 0..63, 64..127
            const int bi = hit ? 64 * q + (int)__builtin_ctzll(hit) : 0x7fffffff;
            if (S.lane == 0) {
                S.red[q] = nan ? __builtin_nan("") : best;
                S.redi[q] = bi;
            }
            if (S.r == bi) S.redv[q] = v;
        }
        __syncthreads();
        double best = S.red[0], akk = S.redv[0];
        int bi = S.redi[0];
        {
            const double ov = S.red[1];
            const int oi = S.redi[1];
            const bool nan = best != best || ov != ov;
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
                akk = S.redv[1];
            }
            if (nan) best = __builtin_nan("");
        }
        bi = __builtin_amdgcn_readfirstlane(bi);
        if (!(best > 0.0)) {
            info = k + 1;
            return false;
        }
        const double inv = 1.0 / akk;
        if (S.r == bi) {
            slot = k;
            if (S.h == HK) {
#pragma unroll
                for (int jl = LB0; jl < half_len<HK>(); jl++) S.pr[HK * kHalf + jl] = row[jl];
                S.dinv[k] = inv;
                S.plane[k] = bi;
            } else if (S.h > HK) {
#pragma unroll
                for (int jl = 0; jl < kHC1; jl++) S.pr[kHalf + jl] = row[jl];
            }
        }
        if (S.h == HK && slot < 0) {
            const double l = sel4<LB0>(row, kk) * inv;
            set4<LB0>(row, kk, l);
            S.lm[S.r] = l;
        }
        __syncthreads();
        if (slot < 0) {
            if (S.h == HK) {
                const double l = sel4<LB0>(row, kk);
                // 8 columns at a time: the scheduling barriers keep the pivot row's loads from
                // all being hoisted ahead of the FMAs (which would take 2 registers per column)
#pragma unroll
                for (int jl = LB0; jl < half_len<HK>(); jl++) {
                    const double lj = (jl < LB0 + kRowChunk && jl + HK * kHalf <= k) ? 0.0 : l;
                    row[jl] = fma(-lj, S.pr[HK * kHalf + jl], row[jl]);
                    if ((jl - LB0) % 8 == 7) __builtin_amdgcn_sched_barrier(0);
                }
            } else if (S.h > HK) {
                const double l = S.lm[S.r];
#pragma unroll
                for (int jl = 0; jl < kHC1; jl++) {
                    row[jl] = fma(-l, S.pr[kHalf + jl], row[jl]);
                    if (jl % 8 == 7) __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
    }
    if constexpr (CH + 1 < kHBlocks) return hr_factor<CH + 1>(row, slot, S, info);
    return true;
}

// back substitution, steps 4 CH + 3 .. 4 CH (blocks descending), in the half holding the block's
// columns: the pivot row of step k turns its sums into x_k = acc / u_kk and publishes it; the rows
// that were pivots before k subtract u_{slot,k} x_k.  Entering half 0, half 1 hands acc over.
template <int CH>
__device__ __forceinline__ void hr_back(const double (&row)[kHN], int slot, double (&acc)[kRhs],
                                        const HalfSys &S, double *xout)
{
    constexpr int B0 = kRowChunk * CH;
    constexpr int HK = B0 >= kHalf ? 1 : 0, LB0 = B0 - HK * kHalf;
    if constexpr (CH == kHSwitch - 1) {
        if (S.h == 1)
#pragma unroll
            for (int c = 0; c < kRhs; c++) S.xfer[S.r * kRhs + c] = acc[c];
        __syncthreads();
        if (S.h == 0)
#pragma unroll
            for (int c = 0; c < kRhs; c++) acc[c] = S.xfer[S.r * kRhs + c];
    }
    for (int kk = kRowChunk - 1; kk >= 0; kk--) {
        const int k = B0 + kk;
        if (k >= S.m) continue;
        if (S.h == HK && slot == k) {
            const double d = S.dinv[k];
#pragma unroll
            for (int c = 0; c < kRhs; c++) xout[k * kRhs + c] = acc[c] * d;
        }
        __syncthreads();
        if (S.h == HK && slot >= 0 && slot < k) {
            const double u = sel4<LB0>(row, kk);
#pragma unroll
            for (int c = 0; c < kRhs; c++) acc[c] = fma(-u, xout[k * kRhs + c], acc[c]);
        }
    }
    if constexpr (CH > 0) hr_back<CH - 1>(row, slot, acc, S, xout);
}

// forward substitution of a new right-hand side with the stored multipliers, in the order of the
// elimination, in the half holding the block's columns; entering half 1, half 0 hands res over
template <int CH>
__device__ __forceinline__ void hr_forward(const double (&row)[kHN], int slot, double (&res)[kRhs],
                                           const HalfSys &S)
{
    constexpr int B0 = kRowChunk * CH;
    constexpr int HK = B0 >= kHalf ? 1 : 0, LB0 = B0 - HK * kHalf;
    if constexpr (CH == kHSwitch) {
        if (S.h == 0)
#pragma unroll
            for (int c = 0; c < kRhs; c++) S.xfer[S.r * kRhs + c] = res[c];
        __syncthreads();
        if (S.h == 1)
#pragma unroll
            for (int c = 0; c < kRhs; c++) res[c] = S.xfer[S.r * kRhs + c];
    }
    for (int kk = 0; kk < kRowChunk; kk++) {
        const int k = B0 + kk;
        if (k >= S.m) return;
        double *b = S.rb + (k & 1) * kRhs;
        if (S.h == HK && slot == k)
#pragma unroll
            for (int c = 0; c < kRhs; c++) b[c] = res[c];
        __syncthreads();
        if (S.h == HK && slot > k && slot < kHMaxA) {
            const double l = sel4<LB0>(row, kk);
#pragma unroll
            for (int c = 0; c < kRhs; c++) res[c] = fma(-l, b[c], res[c]);
        }
    }
    if constexpr (CH + 1 < kHBlocks) hr_forward<CH + 1>(row, slot, res, S);
}

// columns 8 cb .. 8 cb + 7 of the half row from the thread's slots of the assembly tile, by
// selects on cb so that every register index stays static
__device__ __forceinline__ void hr_load_block(double (&row)[kHN], int cb, const double *tile, int tid)
{
    double t8[8];
#pragma unroll
    for (int t = 0; t < 8; t++) t8[t] = tile[t * 256 + tid];
#pragma unroll
    for (int b = 0; b < kHN / 8; b++)
#pragma unroll
        for (int t = 0; t < 8; t++) row[8 * b + t] = cb == b ? t8[t] : row[8 * b + t];
}

// kind 0: the gradient systems; kind 1: the Delta^k systems.  Every work item's system has at
// most 112 unknowns and the back substitution's hand-over needs m > 56.
// 2 wavefronts per SIMD (two systems per CU): the 256 registers then spill about 380 bytes per
// lane to scratch, and still the config-C solves take 298 ms against 348 ms at one wavefront per
// SIMD without spills (every step's barriers and broadcasts wait less with a second system to
// switch to); the weights are bit for bit the same.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
rowhalf_kernel(WArgs a, int kind)
{
    __shared__ __align__(16) double X[3 * kHMaxA];   // stencil coordinates in the solve frame
    __shared__ __align__(16) double pr[kHMaxA + kRhs];   // pivot row (columns of [A | B])
    __shared__ double lm[128];                       // multipliers of the step, by row
    __shared__ __align__(16) double xs[kHMaxA * kRhs];   // solution
    __shared__ __align__(16) double xd[kHMaxA * kRhs];   // refinement correction
    __shared__ __align__(16) double xfer[128 * kRhs];    // half-to-half hand-over, residual halves
    __shared__ double dinv[kHMaxA];
    __shared__ __align__(16) double rb[2 * kRhs];
    __shared__ double red[2], redv[2];
    __shared__ int redi[2];
    __shared__ double tile[8 * 256];                 // assembly: 8 columns per thread
    __shared__ int plane[kHMaxA];
    __shared__ int item_s;
    __shared__ double scale_s;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = tid & 127, h = tid >> 7;
    const int stride = a.n;
    const bool gauss_hv = a.ref_mode || a.hvphs == 0;
    for (;;) {
        if (tid == 0) item_s = atomicAdd(a.counter, 1);
        __syncthreads();
        const int item = item_s;
        __syncthreads();
        if (item >= a.count) break;
        const int node = a.order ? a.order[item] : item;
        const int n = a.nn ? a.nn[node] : stride;
        const int deg = a.dg ? a.dg[node] : a.deg;
        const int M = a.ref_mode ? 21 : (deg + 1) * (deg + 2) / 2;
        const int m = kind == 0 ? n + M : (gauss_hv ? n + 1 : n + M);
        double scale = 1.0;
        if (tid < 64) {
            scale = stencil_frame(a, node, n, X, kHMaxA, lane);
            if (lane == 0) scale_s = scale;
        }
        __syncthreads();
        scale = scale_s;
        WArgs ag = a;
        if (kind == 1) ag.eps = a.eps * scale * (a.nn && n < stride ? sqrt((double)stride / n) : 1.0);
        // ---- assemble the half row: A's entries of this half, then (half 1) the right-hand sides ----
        double row[kHN];
#pragma unroll
        for (int jl = 0; jl < kHN; jl++) row[jl] = 0.0;
        const int j0 = h * kHalf;
        const int alen = h == 0 ? kHalf : kHA1;
        // each thread reads back only its own tile slots: no barrier
        for (int cb = 0; 8 * cb < alen; cb++) {
            for (int t = 0; t < 8; t++) {
                const int jl = 8 * cb + t, j = j0 + jl;
                tile[t * 256 + tid] = (r < m && jl < alen && j < m) ? entry(r, j, n, m, kind, ag, X, kHMaxA) : 0.0;
            }
            hr_load_block(row, cb, tile, tid);
        }
        if (h == 1 && r < m) {
            double b[kRhs];
            rhs_of(r, n, kind, ag, X, kHMaxA, b);
#pragma unroll
            for (int c = 0; c < kRhs; c++) row[kHA1 + c] = b[c];
        }
        int slot = r < m ? -1 : kHMaxA;
        HalfSys S{pr, lm, dinv, red, redv, xfer, rb, plane, redi, r, h, lane, wv, m};
        int info = 0;
        hr_factor<0>(row, slot, S, info);
        __syncthreads();
        if (!info) {
            double acc[kRhs];
#pragma unroll
            for (int c = 0; c < kRhs; c++) acc[c] = h == 1 ? row[kHA1 + c] : 0.0;
            hr_back<kHBlocks - 1>(row, slot, acc, S, xs);
            __syncthreads();
            // ---- one refinement step: r = b - A x (A re-evaluated, each half its columns) ----
            double res[kRhs];
#pragma unroll
            for (int c = 0; c < kRhs; c++) res[c] = 0.0;
            if (r < m) {
                if (h == 0) rhs_of(r, n, kind, ag, X, kHMaxA, res);
                for (int jl = 0; jl < alen; jl++) {
                    const int j = j0 + jl;
                    if (j >= m) break;
                    const double aij = entry(r, j, n, m, kind, ag, X, kHMaxA);
#pragma unroll
                    for (int c = 0; c < kRhs; c++) res[c] = fma(-aij, xs[j * kRhs + c], res[c]);
                }
            }
            if (h == 1)
#pragma unroll
                for (int c = 0; c < kRhs; c++) xfer[r * kRhs + c] = res[c];
            __syncthreads();
            if (h == 0)
#pragma unroll
                for (int c = 0; c < kRhs; c++) res[c] += xfer[r * kRhs + c];
            __syncthreads();
            hr_forward<0>(row, slot, res, S);
            __syncthreads();
            hr_back<kHBlocks - 1>(row, slot, res, S, xd);
            __syncthreads();
        }
        // ---- weights ----
        const size_t rowo = (size_t)node * stride;
        if (kind == 0) {
            for (int s = tid; s < stride; s += 256) {
                const bool in = s < n && !info;
                a.Wx[rowo + s] = in ? (xs[s * kRhs] + xd[s * kRhs]) / scale : 0.0;
                a.Wy[rowo + s] = in ? (xs[s * kRhs + 1] + xd[s * kRhs + 1]) / scale : 0.0;
            }
            if (a.WL1) {
                const double s2 = scale * scale, s4 = s2 * s2;
                for (int s = tid; s < stride; s += 256) {
                    const bool in = s < n && !info;
                    a.WL1[rowo + s] = in ? (xs[s * kRhs + 2] + xd[s * kRhs + 2]) / s2 : 0.0;
                    a.WL2[rowo + s] = in ? (xs[s * kRhs + 3] + xd[s * kRhs + 3]) / s4 : 0.0;
                }
            }
            if (tid == 0) a.status[node] = info ? 1 : 0;
        } else {
            const double s2k = pow(scale, 2.0 * a.hv);
            for (int s = tid; s < stride; s += 256)
                a.WL[rowo + s] = (s < n && !info) ? (xs[s * kRhs] + xd[s * kRhs]) / s2k : 0.0;
            if (tid == 0 && info && a.status[node] == 0) a.status[node] = 2;
        }
        __syncthreads();
    }
}

}  // namespace

// systems of 65 .. 112 unknowns (the class is used above 96)
int launch_rowhalf(serbf_engine *e, const WArgs &a, int kind)
{
    const void *fn = (const void *)rowhalf_kernel;
    if (e->cus <= 0) {
        hipDeviceProp_t prop;
        e->cus = hipGetDeviceProperties(&prop, e->device) == hipSuccess ? prop.multiProcessorCount : 256;
    }
    // resident work-groups only (the kernel is persistent)
    int per_cu = 0;
    SERBF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0));
    if (per_cu < 1) per_cu = 1;
    int grid = e->cus * per_cu;
    if (grid > a.count) grid = a.count;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(rowhalf_kernel, dim3(grid), dim3(256), 0, e->stream, a, kind);
    SERBF_HIP(hipGetLastError());
    return SERBF_OK;
}

}  // namespace serbf
