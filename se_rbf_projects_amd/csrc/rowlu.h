// Register-row batched LU of the per-node weight systems (see weights.hip for the LDS
// kernel and the schedule); each class is instantiated in its own translation unit
// (rowlu_*.hip), which keeps the build parallel.
#pragma once
#include "weights_common.h"

namespace serbf {

// ---------------------------------------------------------------------------
// Register-row batched LU: one work-group of W wavefronts per system (W = 1 up to 64 rows, 2
// up to 128), thread t holding row t of [A | B] in registers: A padded with zeros to MAXA
// columns, then the four right-hand-side columns.  Per step k of the elimination:
//   - the work-group picks the pivot among the rows not yet used (largest |a_ik|, the lower
//     row index on ties): the wavefront maximum by DPP steps (quad, half-row, row, row
//     broadcasts: no LDS round trip), the lowest lane holding it by a ballot, then across the
//     W wavefronts through LDS;
//   - with W = 2 the pivot's thread writes its row (columns from the current block of 4 on) to
//     LDS and every thread reads it back as one broadcast; with W = 1 every lane reads it
//     straight from the pivot lane's registers (v_readlane), so a step has no LDS access and
//     no barrier (the substitutions likewise).  Against the shuffle butterfly and the LDS row
//     for every class this took the 655362-node weights from 252 to 222 ms, bit for bit the
//     same results;
//   - every row not yet used stores its multiplier l = a_ik / a_kk in column k and updates
//     columns k+1 .. MAXC-1 with register FMAs.
// Rows are not swapped: the step at which a row became the pivot (its slot) is kept, and the
// pivot row of step k (plane[k]) is U's row k with its right-hand side y_k.  Column k of a
// register row is a run-time index inside a block of 4 columns and a compile-time one
// outside it, so each block of 4 steps is its own straight-line code with static register
// indices (template recursion over the blocks).  Back substitution, the refinement residual
// (A re-evaluated from the coordinates) and its solve with the same factors follow.  The
// matrix is assembled 8 columns at a time through an LDS tile.  Compared with the LDS
// kernel above (rows in LDS, one LDS load and store per multiply-add), the multiply-adds read
// and write registers only and the LDS holds just the coordinates, the pivot row, the
// solution and the tile (8-16 KB per system), so many more systems run per CU.  A row of
// MAXA + 4 doubles must fit the 256 addressable vector registers with room to spare: the
// classes stop at MAXA = 96.
constexpr int kRowChunk = 4;

template <int W, int MAXA>
struct RowLds {
    static constexpr int MAXC = MAXA + kRhs;
    // X[3][MAXA] | pr[MAXC] | xs[MAXA][kRhs] | xd[MAXA][kRhs] | dinv[MAXA] | rb[2][kRhs] | red[2 W] |
    // tile[8][64 W] | plane[MAXA] | redi[2 W] (int)
    static constexpr int doubles = 3 * MAXA + MAXC + 2 * MAXA * kRhs + MAXA + 2 * kRhs + 2 * W + 8 * 64 * W;
    static constexpr int bytes = doubles * 8 + (MAXA + 2 * W) * 4;
};

// v[B0 + kk] for a run-time kk < 4 (B0 static)
template <int B0, int N>
__device__ __forceinline__ double sel4(const double (&v)[N], int kk)
{
    return kk == 0 ? v[B0] : kk == 1 ? v[B0 + 1] : kk == 2 ? v[B0 + 2] : v[B0 + 3];
}
// v[B0 + kk] = x for a run-time kk < 4, as selects (a conditional store here can be merged
// into one store at a run-time index, which would move the row out of registers)
template <int B0, int N>
__device__ __forceinline__ void set4(double (&v)[N], int kk, double x)
{
#pragma unroll
    for (int t = 0; t < kRowChunk; t++) v[B0 + t] = kk == t ? x : v[B0 + t];
}

// columns 8 cb .. 8 cb + 7 of the tile into the register row: the tile's 8 values are loaded
// once and every block of the row takes them by a select on cb, so every register index stays
// static (a branch per block would be merged into one store through a run-time address, which
// moves the row out of registers)
template <int W, int MAXA>
__device__ __forceinline__ void rl_load_block(double (&row)[MAXA + kRhs], int cb, const double *tile, int tid)
{
    double t8[8];
#pragma unroll
    for (int t = 0; t < 8; t++) t8[t] = tile[t * 64 * W + tid];
#pragma unroll
    for (int b = 0; b < MAXA / 8; b++)
#pragma unroll
        for (int t = 0; t < 8; t++) row[8 * b + t] = cb == b ? t8[t] : row[8 * b + t];
}

// lane src's value of x, in every lane (src wave-uniform)
__device__ __forceinline__ double bcast(double x, int src)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// one DPP step of a 64-bit value: lanes inside the row mask take the value of the lane the
// control names, the others keep their own
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp64(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
    const int l2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xf, false);
    const int h2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xf, false);
    return __longlong_as_double(((long long)h2 << 32) | (unsigned)l2);
}

// largest of v over the wavefront (NaN ignored), in every lane: quad, half-row and row steps,
// then the row broadcasts carry rows 0-2 into lane 63
__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp64<0xB1, 0xf>(v));    // quad_perm [1, 0, 3, 2]
    v = fmax(v, dpp64<0x4E, 0xf>(v));    // quad_perm [2, 3, 0, 1]
    v = fmax(v, dpp64<0x141, 0xf>(v));   // row_half_mirror
    v = fmax(v, dpp64<0x140, 0xf>(v));   // row_mirror
    v = fmax(v, dpp64<0x142, 0xa>(v));   // row_bcast:15 into rows 1 and 3
    v = fmax(v, dpp64<0x143, 0xc>(v));   // row_bcast:31 into rows 2 and 3
    return bcast(v, 63);
}

struct RowSys {
    double *pr, *dinv, *red, *rb;
    int *plane, *redi;
    int tid, m;
};

// steps 4 CH .. 4 CH + 3 of the elimination; false once k reaches m or a pivot is zero
template <int W, int MAXA, int CH>
__device__ __forceinline__ bool rl_factor(double (&row)[MAXA + kRhs], int &slot, const RowSys &S, int &info)
{
    constexpr int MAXC = MAXA + kRhs, B0 = kRowChunk * CH;
    const int tid = S.tid, lane = tid & 63, w = tid >> 6;
    for (int kk = 0; kk < kRowChunk; kk++) {
        const int k = B0 + kk;
        if (k >= S.m) return false;
        // the pivot: the largest |a_ik| among the rows not yet used, the lowest row on ties; a NaN
        // candidate marks the system singular
        const double mine = slot < 0 ? fabs(sel4<B0>(row, kk)) : -1.0;
        double best = wave_max(mine);
        const unsigned long long hit = __ballot(slot < 0 && mine == best);
        int bi = hit ? 64 * w + (int)__builtin_ctzll(hit) : 0x7fffffff;
        if (__ballot(slot < 0 && mine != mine)) best = __builtin_nan("");
        if (W > 1) {
            double *red = S.red + (k & 1) * W;
            int *redi = S.redi + (k & 1) * W;
            if (lane == 0) {
                red[w] = best;
                redi[w] = bi;
            }
            __syncthreads();
            best = red[0];
            bi = redi[0];
            bool nan = best != best;
#pragma unroll
            for (int v = 1; v < W; v++) {
                const double ov = red[v];
                const int oi = redi[v];
                nan |= ov != ov;
                if (ov > best || (ov == best && oi < bi)) {
                    best = ov;
                    bi = oi;
                }
            }
            if (nan) best = __builtin_nan("");
        }
        bi = __builtin_amdgcn_readfirstlane(bi);
        if (!(best > 0.0)) {
            info = k + 1;
            return false;
        }
        if constexpr (W == 1) {
            // one wavefront: the pivot row is read from its lane's registers, no LDS, no barrier
            const double inv = 1.0 / bcast(sel4<B0>(row, kk), bi);
            if (tid == bi) {
                slot = k;
                S.dinv[k] = inv;
                S.plane[k] = bi;
            }
            double l = 0.0;
            if (slot < 0) {
                l = sel4<B0>(row, kk) * inv;
                set4<B0>(row, kk, l);
            }
#pragma unroll
            for (int j = B0; j < MAXC; j++) {
                const double lj = (j < B0 + kRowChunk && j <= k) ? 0.0 : l;
                row[j] = fma(-lj, bcast(row[j], bi), row[j]);
            }
        } else {
        if (tid == bi) {
#pragma unroll
            for (int j = B0; j < MAXC; j++) S.pr[j] = row[j];
            slot = k;
            S.dinv[k] = 1.0 / sel4<B0>(row, kk);
            S.plane[k] = bi;
        }
        __syncthreads();
        const double inv = S.dinv[k];
        double l = 0.0;
        if (slot < 0) {
            l = sel4<B0>(row, kk) * inv;
            set4<B0>(row, kk, l);
        }
#pragma unroll
        for (int j = B0; j < MAXC; j++) {
            const double lj = (j < B0 + kRowChunk && j <= k) ? 0.0 : l;
            row[j] = fma(-lj, S.pr[j], row[j]);
        }
        __syncthreads();
        }
    }
    if constexpr (CH + 1 < MAXA / kRowChunk) return rl_factor<W, MAXA, CH + 1>(row, slot, S, info);
    return true;
}

// back substitution, steps 4 CH + 3 .. 4 CH (blocks in descending order): the pivot row of
// step k turns its accumulated right-hand side into x_k = acc / u_kk and publishes it; the rows
// that were pivots before k subtract u_{slot,k} x_k
template <int W, int MAXA, int CH>
__device__ __forceinline__ void rl_back(const double (&row)[MAXA + kRhs], int slot, double (&acc)[kRhs],
                                        const RowSys &S, double *xout)
{
    constexpr int B0 = kRowChunk * CH;
    for (int kk = kRowChunk - 1; kk >= 0; kk--) {
        const int k = B0 + kk;
        if (k >= S.m) continue;
        if constexpr (W == 1) {
            const int pk = __builtin_amdgcn_readfirstlane(S.plane[k]);
            const double d = S.dinv[k];
            double x[kRhs];
#pragma unroll
            for (int c = 0; c < kRhs; c++) x[c] = bcast(acc[c], pk) * d;
            if (slot == k)
#pragma unroll
                for (int c = 0; c < kRhs; c++) xout[k * kRhs + c] = x[c];
            if (slot >= 0 && slot < k) {
                const double u = sel4<B0>(row, kk);
#pragma unroll
                for (int c = 0; c < kRhs; c++) acc[c] = fma(-u, x[c], acc[c]);
            }
            continue;
        }
        if (slot == k) {
            const double d = S.dinv[k];
#pragma unroll
            for (int c = 0; c < kRhs; c++) xout[k * kRhs + c] = acc[c] * d;
        }
        __syncthreads();
        if (slot >= 0 && slot < k) {
            const double u = sel4<B0>(row, kk);
#pragma unroll
            for (int c = 0; c < kRhs; c++) acc[c] = fma(-u, xout[k * kRhs + c], acc[c]);
        }
    }
    if constexpr (CH > 0) rl_back<W, MAXA, CH - 1>(row, slot, acc, S, xout);
}

// forward substitution of a new right-hand side with the stored multipliers, in the order of
// the elimination: the pivot row of step k publishes its value, the rows used later subtract
// l_{row,k} times it
template <int W, int MAXA, int CH>
__device__ __forceinline__ void rl_forward(const double (&row)[MAXA + kRhs], int slot, double (&res)[kRhs],
                                           const RowSys &S)
{
    constexpr int B0 = kRowChunk * CH;
    for (int kk = 0; kk < kRowChunk; kk++) {
        const int k = B0 + kk;
        if (k >= S.m) return;
        if constexpr (W == 1) {
            const int pk = __builtin_amdgcn_readfirstlane(S.plane[k]);
            double bv[kRhs];
#pragma unroll
            for (int c = 0; c < kRhs; c++) bv[c] = bcast(res[c], pk);
            if (slot > k && slot < MAXA) {
                const double l = sel4<B0>(row, kk);
#pragma unroll
                for (int c = 0; c < kRhs; c++) res[c] = fma(-l, bv[c], res[c]);
            }
            continue;
        }
        double *b = S.rb + (k & 1) * kRhs;
        if (slot == k)
#pragma unroll
            for (int c = 0; c < kRhs; c++) b[c] = res[c];
        __syncthreads();
        if (slot > k && slot < MAXA) {
            const double l = sel4<B0>(row, kk);
#pragma unroll
            for (int c = 0; c < kRhs; c++) res[c] = fma(-l, b[c], res[c]);
        }
    }
    if constexpr (CH + 1 < MAXA / kRowChunk) rl_forward<W, MAXA, CH + 1>(row, slot, res, S);
}

// kind 0: the gradient systems; kind 1: the Delta^k systems.  Every work item's system fits
// the class (m <= MAXA <= 64 W, n <= MAXA).
// Occupancy target per class: 2 wavefronts per SIMD (256 registers), which holds every class's
// row of MAXA + 4 doubles and the rest without spilling to scratch (at 3 wavefronts per SIMD
// the 48-column class spilled, and its results depended on where the spills fell).
template <int MAXA>
constexpr int kRowWaves = 2;

template <int W, int MAXA>
__global__ void __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(kRowWaves<MAXA>)))
rowlu_kernel(WArgs a, int kind)
{
    constexpr int MAXC = RowLds<W, MAXA>::MAXC;
    // static LDS arrays: their addresses are constants, so each access is one LDS instruction
    // with an immediate offset
    __shared__ __align__(16) double X[3 * MAXA];               // stencil coordinates in the solve frame
    __shared__ __align__(16) double pr[MAXC];                  // pivot row of the current step
    __shared__ __align__(16) double xs[MAXA * kRhs];           // solution
    __shared__ __align__(16) double xd[MAXA * kRhs];           // refinement correction
    __shared__ double dinv[MAXA];                // 1 / u_kk
    __shared__ __align__(16) double rb[2 * kRhs];
    __shared__ double red[2 * W];                // per-wavefront pivot candidates
    __shared__ double tile[8 * 64 * W];          // assembly tile
    __shared__ int plane[MAXA];                  // pivot row of each step
    __shared__ int redi[2 * W];
    __shared__ int item_s;
    const int tid = threadIdx.x, lane = tid & 63;
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
        // the stencil frame: wavefront 0 builds it, the others wait
        double scale = 1.0;
        if (tid < 64) {
            scale = stencil_frame(a, node, n, X, MAXA, lane);
            if (lane == 0) red[0] = scale;
        }
        __syncthreads();
        scale = red[0];
        __syncthreads();
        WArgs ag = a;
        if (kind == 1) ag.eps = a.eps * scale * (a.nn && n < stride ? sqrt((double)stride / n) : 1.0);
        // ---- assemble [A | B] ----
        double row[MAXC];
#pragma unroll
        for (int j = 0; j < MAXC; j++) row[j] = 0.0;
        int slot = tid < m ? -1 : MAXA;
        if (tid < m) {
            double b[kRhs];
            rhs_of(tid, n, kind, ag, X, MAXA, b);
#pragma unroll
            for (int c = 0; c < kRhs; c++) row[MAXA + c] = b[c];
        }
        for (int cb = 0; cb * 8 < m; cb++) {
            for (int t = 0; t < 8; t++) {
                const int j = 8 * cb + t;
                tile[t * 64 * W + tid] = (tid < m && j < m) ? entry(tid, j, n, m, kind, ag, X, MAXA) : 0.0;
            }
            __syncthreads();
            rl_load_block<W, MAXA>(row, cb, tile, tid);
            __syncthreads();
        }
        // ---- factor, solve ----
        RowSys S{pr, dinv, red, rb, plane, redi, tid, m};
        int info = 0;
        rl_factor<W, MAXA, 0>(row, slot, S, info);
        __syncthreads();
        if (!info) {
            double acc[kRhs];
#pragma unroll
            for (int c = 0; c < kRhs; c++) acc[c] = row[MAXA + c];
            rl_back<W, MAXA, MAXA / kRowChunk - 1>(row, slot, acc, S, xs);
            __syncthreads();
            // ---- one refinement step: r = b - A x (A re-evaluated), solved with the factors ----
#pragma unroll
            for (int c = 0; c < kRhs; c++) acc[c] = 0.0;
            if (tid < m) {
                double b[kRhs];
                rhs_of(tid, n, kind, ag, X, MAXA, b);
#pragma unroll
                for (int c = 0; c < kRhs; c++) acc[c] = b[c];
                for (int j = 0; j < m; j++) {
                    const double aij = entry(tid, j, n, m, kind, ag, X, MAXA);
#pragma unroll
                    for (int c = 0; c < kRhs; c++) acc[c] = fma(-aij, xs[j * kRhs + c], acc[c]);
                }
            }
            rl_forward<W, MAXA, 0>(row, slot, acc, S);
            __syncthreads();
            rl_back<W, MAXA, MAXA / kRowChunk - 1>(row, slot, acc, S, xd);
            __syncthreads();
        }
        // ---- weights ----
        const size_t rowo = (size_t)node * stride;
        if (kind == 0) {
            for (int s = tid; s < stride; s += 64 * W) {
                const bool in = s < n && !info;
                a.Wx[rowo + s] = in ? (xs[s * kRhs] + xd[s * kRhs]) / scale : 0.0;
                a.Wy[rowo + s] = in ? (xs[s * kRhs + 1] + xd[s * kRhs + 1]) / scale : 0.0;
            }
            if (a.WL1) {
                const double s2 = scale * scale, s4 = s2 * s2;
                for (int s = tid; s < stride; s += 64 * W) {
                    const bool in = s < n && !info;
                    a.WL1[rowo + s] = in ? (xs[s * kRhs + 2] + xd[s * kRhs + 2]) / s2 : 0.0;
                    a.WL2[rowo + s] = in ? (xs[s * kRhs + 3] + xd[s * kRhs + 3]) / s4 : 0.0;
                }
            }
            if (tid == 0) a.status[node] = info ? 1 : 0;
        } else {
            const double s2k = pow(scale, 2.0 * a.hv);
            for (int s = tid; s < stride; s += 64 * W)
                a.WL[rowo + s] = (s < n && !info) ? (xs[s * kRhs] + xd[s * kRhs]) / s2k : 0.0;
            if (tid == 0 && info && a.status[node] == 0) a.status[node] = 2;
        }
        __syncthreads();
    }
}

template <int W, int MAXA>
int launch_rowlu_t(serbf_engine *e, const WArgs &a, int kind)
{
    const int bytes = 0;   // static LDS (RowLds::bytes)
    const void *fn = (const void *)rowlu_kernel<W, MAXA>;
    if (e->cus <= 0) {
        hipDeviceProp_t prop;
        e->cus = hipGetDeviceProperties(&prop, e->device) == hipSuccess ? prop.multiProcessorCount : 256;
    }
    // resident work-groups only (the kernel is persistent)
    int per_cu = 0;
    SERBF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * W, bytes));
    if (per_cu < 1) per_cu = 1;
    int grid = e->cus * per_cu;
    if (grid > a.count) grid = a.count;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL((rowlu_kernel<W, MAXA>), dim3(grid), dim3(64 * W), bytes, e->stream, a, kind);
    SERBF_HIP(hipGetLastError());
    return SERBF_OK;
}

int launch_rowlu_48(serbf_engine *e, const WArgs &a, int kind);
int launch_rowlu_56(serbf_engine *e, const WArgs &a, int kind);
int launch_rowlu_64(serbf_engine *e, const WArgs &a, int kind);
int launch_rowlu_72(serbf_engine *e, const WArgs &a, int kind);
int launch_rowlu_80(serbf_engine *e, const WArgs &a, int kind);
int launch_rowlu_88(serbf_engine *e, const WArgs &a, int kind);
int launch_rowlu_96(serbf_engine *e, const WArgs &a, int kind);
// half-row kernel (rowhalf.hip): systems of 65 .. 112 unknowns, two threads per row
int launch_rowhalf(serbf_engine *e, const WArgs &a, int kind);

}  // namespace serbf
