// Batched per-node RBF-FD weight solves (gfx950), the body of get_rbffd_DMs
// (rbffd_DMs.c:98-282 of the reference).
//
// Persistent launches factor every node's saddle systems; work is pulled from a global counter
// so uneven system sizes balance across the chip.  Two kernels:
//   - the register-row kernel (rowlu.h, the default schedule): one work-group per system,
//     one row of [A | B] per thread in registers; one launch per size class and system kind;
//   - the LDS kernel below: one wavefront per node, which assembles [Phi P; P^T 0] and the
//     right-hand sides in LDS (lane = row, rows lane + 64 r), factors it with partial pivoting
//     (first maximal |a_ik|, full-row swaps as LAPACK dgesv) and solves both of the node's
//     systems; it takes the systems larger than the register classes.
// Both apply one step of iterative refinement whose residual b - A x re-evaluates A from the
// coordinates (the matrix itself is never stored in HBM).
//
// Systems per node:
//   reference mode: PHS r^k (k = phs_order) + degree-5 tangent monomials, RHS
//     d/dx', d/dy' of r^k at the centre (rbffd_DMs.c:175-215, 240); Gaussian +
//     constant with RHS Delta^k by the Laguerre recursion (rbffd_DMs.c:188,198,
//     223-228, 243, 314-328); coordinates unscaled.
//   config mode: the stencil is shifted to the centre and scaled by its radius s;
//     PHS r^k + degree-d monomials for the gradient (weights / s) and PHS r^K +
//     the same monomials with RHS Delta^k r^K = prod (K-2j)^2 r^(K-2k) for the
//     hyperviscosity operator (weights / s^2k).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rowlu.h"
#include "weights_common.h"

namespace serbf {

// LDS layout of one wavefront's system (doubles):
//   A   [MAXM][LD]  augmented matrix [A | B] (LD >= MAXM + 4, odd)
//   X   [3][MAXN]   stencil coordinates in the solve frame
//   B0  [MAXM][4]   right-hand sides (kept for the residual)
//   XS  [MAXM][4]   solution
//   piv [MAXM]      pivot rows (int)

//   A0  [MAXM][MAXM] (COPY variant) the assembled matrix, for the refinement residual
template <int MAXM, bool COPY>
struct Lds {
    static constexpr int LD = ((MAXM + 4) | 1);
    static constexpr int MAXN = MAXM;
    static constexpr int doubles = MAXM * LD + 3 * MAXN + 2 * MAXM * kRhs + (COPY ? MAXM * MAXM : 0);
    static constexpr int bytes = doubles * 8 + MAXM * 4;
};

template <int MAXM, bool COPY>
__device__ int factor_solve(double *A, double *X, double *B0, double *XS, const double *A0, int *piv, int n, int m,
                            int nrhs, int kind, const WArgs &a, int lane)
{
    constexpr int LD = Lds<MAXM, COPY>::LD, MAXN = Lds<MAXM, COPY>::MAXN, R = (MAXM + 63) / 64;
    int info = 0;
    // ---- LU with partial pivoting on [A | B] ----
    for (int k = 0; k < m; k++) {
        double best = -1.0;
        int bi = m;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = lane + 64 * r;
            if (i >= k && i < m) {
                const double v = fabs(A[i * LD + k]);
                if (v > best) { best = v; bi = i; }
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        const int p = __builtin_amdgcn_readfirstlane(bi);
        if (lane == 0) piv[k] = p;
        if (!(best > 0.0)) {
            if (!info) info = k + 1;
            wave_sync();
            continue;
        }
        if (p != k)
            for (int j = lane; j < m + nrhs; j += 64) {
                const double t = A[k * LD + j];
                A[k * LD + j] = A[p * LD + j];
                A[p * LD + j] = t;
            }
        wave_sync();
        const double pv = A[k * LD + k];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int i = lane + 64 * r;
            if (i > k && i < m) {
                const double l = A[i * LD + k] / pv;
                A[i * LD + k] = l;
                if (l != 0.0) {
                    // groups of 4 independent loads before the stores: the row update is a
                    // latency chain through LDS otherwise
                    double *ri = A + i * LD;
                    const double *rk = A + k * LD;
                    const int mm = m + nrhs;
                    int j = k + 1;
                    for (; j + 4 <= mm; j += 4) {
                        const double a0 = ri[j], a1 = ri[j + 1], a2 = ri[j + 2], a3 = ri[j + 3];
                        const double b0 = rk[j], b1 = rk[j + 1], b2 = rk[j + 2], b3 = rk[j + 3];
                        ri[j] = a0 - l * b0;
                        ri[j + 1] = a1 - l * b1;
                        ri[j + 2] = a2 - l * b2;
                        ri[j + 3] = a3 - l * b3;
                    }
                    for (; j < mm; j++) ri[j] -= l * rk[j];
                }
            }
        }
        wave_sync();
    }
    if (info) return info;
    // ---- back substitution U x = y (y in the augmented columns) ----
    for (int i = m - 1; i >= 0; i--) {
        if (lane < nrhs) A[i * LD + m + lane] /= A[i * LD + i];
        wave_sync();
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = lane + 64 * r;
            if (row < i) {
                const double u = A[row * LD + i];
                for (int c = 0; c < nrhs; c++) A[row * LD + m + c] -= u * A[i * LD + m + c];
            }
        }
        wave_sync();
    }
    for (int i = lane; i < m; i += 64)
        for (int c = 0; c < nrhs; c++) XS[i * kRhs + c] = A[i * LD + m + c];
    wave_sync();
    // ---- one refinement step: r = b - A x with A re-evaluated, solve with the factors ----
    double res[R][kRhs];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = lane + 64 * r;
        for (int c = 0; c < kRhs; c++) res[r][c] = 0.0;
        if (i < m) {
            double acc[kRhs];
            for (int c = 0; c < kRhs; c++) acc[c] = B0[i * kRhs + c];
            for (int j = 0; j < m; j++) {
                const double aij = COPY ? A0[i * MAXM + j] : entry(i, j, n, m, kind, a, X, MAXN);
                for (int c = 0; c < nrhs; c++) acc[c] -= aij * XS[j * kRhs + c];
            }
            for (int c = 0; c < nrhs; c++) res[r][c] = acc[c];
        }
    }
    // store the residual in the augmented columns, apply the row swaps, then L and U solves
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = lane + 64 * r;
        if (i < m)
            for (int c = 0; c < nrhs; c++) A[i * LD + m + c] = res[r][c];
    }
    wave_sync();
    if (lane < nrhs)
        for (int k = 0; k < m; k++) {
            const int p = piv[k];
            if (p != k) {
                const double t = A[k * LD + m + lane];
                A[k * LD + m + lane] = A[p * LD + m + lane];
                A[p * LD + m + lane] = t;
            }
        }
    wave_sync();
    for (int k = 0; k < m; k++) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = lane + 64 * r;
            if (row > k && row < m) {
                const double l = A[row * LD + k];
                for (int c = 0; c < nrhs; c++) A[row * LD + m + c] -= l * A[k * LD + m + c];
            }
        }
        wave_sync();
    }
    for (int i = m - 1; i >= 0; i--) {
        if (lane < nrhs) A[i * LD + m + lane] /= A[i * LD + i];
        wave_sync();
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int row = lane + 64 * r;
            if (row < i) {
                const double u = A[row * LD + i];
                for (int c = 0; c < nrhs; c++) A[row * LD + m + c] -= u * A[i * LD + m + c];
            }
        }
        wave_sync();
    }
    for (int i = lane; i < m; i += 64)
        for (int c = 0; c < nrhs; c++) XS[i * kRhs + c] += A[i * LD + m + c];
    wave_sync();
    return 0;
}

template <int MAXM, bool COPY>
__device__ void assemble(double *A, double *X, double *B0, double *A0, int n, int m, int nrhs, int kind,
                         const WArgs &a, int lane)
{
    constexpr int LD = Lds<MAXM, COPY>::LD, MAXN = Lds<MAXM, COPY>::MAXN;
    for (int i = lane; i < m; i += 64) {
        for (int j = 0; j < m; j++) {
            const double v = entry(i, j, n, m, kind, a, X, MAXN);
            A[i * LD + j] = v;
            if (COPY) A0[i * MAXM + j] = v;
        }
        double b[kRhs];
        rhs_of(i, n, kind, a, X, MAXN, b);
        for (int c = 0; c < nrhs; c++) {
            A[i * LD + m + c] = b[c];
            B0[i * kRhs + c] = b[c];
        }
        for (int c = nrhs; c < kRhs; c++) B0[i * kRhs + c] = 0.0;
    }
    wave_sync();
}

template <int MAXM, bool COPY>
__global__ void __launch_bounds__(64) weights_kernel(WArgs a)
{
    extern __shared__ double smem[];
    constexpr int LD = Lds<MAXM, COPY>::LD, MAXN = Lds<MAXM, COPY>::MAXN;
    double *A = smem;
    double *X = A + MAXM * LD;
    double *B0 = X + 3 * MAXN;
    double *XS = B0 + MAXM * kRhs;
    double *A0 = XS + MAXM * kRhs;   // used by the COPY variant only
    int *piv = (int *)(A0 + (COPY ? MAXM * MAXM : 0));
    const int lane = threadIdx.x;
    const int stride = a.n;   // row stride of the stencil and weight arrays
    const bool gauss_hv = a.ref_mode || a.hvphs == 0;
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(a.counter, 1);
        item = __builtin_amdgcn_readfirstlane(__shfl(item, 0, 64));
        if (item >= a.count) break;
        const int node = a.order ? a.order[item] : item;
        const int n = a.nn ? a.nn[node] : stride;
        const int deg = a.dg ? a.dg[node] : a.deg;
        const int M = a.ref_mode ? 21 : (deg + 1) * (deg + 2) / 2;
        const int m1 = n + M;
        const int m2 = gauss_hv ? n + 1 : n + M;
        const double scale = stencil_frame(a, node, n, X, MAXN, lane);
        int st = 0;
        const size_t row = (size_t)node * stride;
        // gradient system (config mode with closure weights: also Delta and Delta^2)
        const int ng = a.WL1 ? 4 : 2;
        assemble<MAXM, COPY>(A, X, B0, A0, n, m1, ng, 0, a, lane);
        if (factor_solve<MAXM, COPY>(A, X, B0, XS, A0, piv, n, m1, ng, 0, a, lane)) st = 1;
        for (int s = lane; s < stride; s += 64) {
            const bool in = s < n && !st;
            a.Wx[row + s] = in ? XS[s * kRhs] / scale : 0.0;
            a.Wy[row + s] = in ? XS[s * kRhs + 1] / scale : 0.0;
        }
        if (a.WL1) {
            const double s2 = scale * scale, s4 = s2 * s2;
            for (int s = lane; s < stride; s += 64) {
                const bool in = s < n && !st;
                a.WL1[row + s] = in ? XS[s * kRhs + 2] / s2 : 0.0;
                a.WL2[row + s] = in ? XS[s * kRhs + 3] / s4 : 0.0;
            }
        }
        wave_sync();
        // Delta^k system
        if (a.hv > 0) {
            // the Gaussian shape parameter of the unit sphere, in the scaled solve frame; with
            // per-node stencil sizes it grows as sqrt(n_max / n), so that eps times the stencil
            // radius is the same for every size
            WArgs ag = a;
            ag.eps = a.eps * scale * (a.nn && n < stride ? sqrt((double)stride / n) : 1.0);
            assemble<MAXM, COPY>(A, X, B0, A0, n, m2, 1, 1, ag, lane);
            int bad = factor_solve<MAXM, COPY>(A, X, B0, XS, A0, piv, n, m2, 1, 1, ag, lane);
            if (bad && !st) st = 2;
            const double s2k = pow(scale, 2.0 * a.hv);
            for (int s = lane; s < stride; s += 64) a.WL[row + s] = (s < n && !bad) ? XS[s * kRhs] / s2k : 0.0;
        } else {
            for (int s = lane; s < stride; s += 64) a.WL[row + s] = 0.0;
        }
        if (lane == 0) a.status[node] = st;
        wave_sync();
    }
}

namespace {

template <int MAXM, bool COPY>
int launch_weights(serbf_engine *e, const WArgs &a, int blocks_per_cu)
{
    const int bytes = Lds<MAXM, COPY>::bytes;
    SERBF_HIP(hipFuncSetAttribute((const void *)weights_kernel<MAXM, COPY>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    if (e->cus <= 0) {
        hipDeviceProp_t prop;
        e->cus = hipGetDeviceProperties(&prop, e->device) == hipSuccess ? prop.multiProcessorCount : 256;
    }
    const int fit = (160 * 1024) / bytes > 0 ? (160 * 1024) / bytes : 1;
    const int per_cu = blocks_per_cu > 0 && blocks_per_cu < fit ? blocks_per_cu : fit;
    int grid = e->cus * per_cu;
    if (grid > a.count) grid = a.count;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL((weights_kernel<MAXM, COPY>), dim3(grid), dim3(64), bytes, e->stream, a);
    SERBF_HIP(hipGetLastError());
    return SERBF_OK;
}

// Schedule variants of the persistent solve: the LDS class (smallest of 48/56/64/72/80/88/112
// rows that holds the largest system: at m = 70 the 72-row class fits 3 wavefronts per CU where
// the 80-row class fits 2), the residual of the refinement step (re-evaluated from the
// coordinates, or read from a copy of the matrix kept in LDS, when it fits), and the work
// order (node order, or largest systems first).
struct WVariant {
    int rows;      // 1: register-row kernel (systems too large for its classes use the LDS kernel)
    int copy;      // 1: matrix copy in LDS (LDS kernel)
    int sorted;    // 1: largest systems first
};
constexpr int kNVariants = 6;
constexpr WVariant kWVariants[kNVariants] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}, {1, 0, 0}, {1, 0, 1}};

// register-row classes: MAXA (columns of A held per row), one wavefront up to 64 rows and two
// above; a row of 96 + 4 doubles is the largest that leaves the kernel room in the 256
// addressable vector registers.  Every 8 columns, so a system carries at most 7 padding
// columns through its multiply-adds and pivot-row reads.
constexpr int kNRowClasses = 7;
constexpr int kRowClasses[kNRowClasses] = {48, 56, 64, 72, 80, 88, 96};

int row_class_of(int m)
{
    for (int ci = 0; ci < kNRowClasses; ci++)
        if (m <= kRowClasses[ci]) return ci;
    return -1;
}

// register schedules of a system of m unknowns: the register-row classes, then the half-row
// class (rowhalf.hip) up to 112 unknowns (index kNRowClasses); -1 none
int reg_class_of(int m)
{
    const int ci = row_class_of(m);
    return ci >= 0 ? ci : (m <= 112 ? kNRowClasses : -1);
}

template <int MAXM>
int class_bytes(int copy)
{
    return copy ? Lds<MAXM, true>::bytes : Lds<MAXM, false>::bytes;
}

bool variant_fits(int maxm, int copy)
{
    const int bytes = maxm <= 48   ? class_bytes<48>(copy)
                      : maxm <= 56 ? class_bytes<56>(copy)
                      : maxm <= 64 ? class_bytes<64>(copy)
                      : maxm <= 72 ? class_bytes<72>(copy)
                      : maxm <= 80 ? class_bytes<80>(copy)
                      : maxm <= 88 ? class_bytes<88>(copy)
                                   : class_bytes<112>(copy);
    return bytes <= 160 * 1024;
}

template <int MAXM>
int launch_in(serbf_engine *e, const WArgs &a, int copy)
{
    return copy ? launch_weights<MAXM, true>(e, a, 0) : launch_weights<MAXM, false>(e, a, 0);
}

int launch_class(serbf_engine *e, const WArgs &a, int maxm, int copy)
{
    if (maxm <= 48) return launch_in<48>(e, a, copy);
    if (maxm <= 56) return launch_in<56>(e, a, copy);
    if (maxm <= 64) return launch_in<64>(e, a, copy);
    if (maxm <= 72) return launch_in<72>(e, a, copy);
    if (maxm <= 80) return launch_in<80>(e, a, copy);
    if (maxm <= 88) return launch_in<88>(e, a, copy);
    return launch_in<112>(e, a, copy);
}

// gradient-system size of every node, the largest one and the order of the nodes by
// decreasing size
int system_sizes(serbf_engine *e, int &maxm, std::vector<int> &by_size, std::vector<int> &m)
{
    const serbf_config &c = e->cfg;
    const int N = e->N;
    m.assign(N, 0);
    if (e->h_nn.empty()) {
        const int M = c.mode == SERBF_MODE_REFERENCE ? 21 : (c.poly_deg + 1) * (c.poly_deg + 2) / 2;
        for (int i = 0; i < N; i++) m[i] = e->n + M;
    } else {
        for (int i = 0; i < N; i++) m[i] = e->h_nn[i] + (e->h_deg[i] + 1) * (e->h_deg[i] + 2) / 2;
    }
    maxm = 0;
    for (int v : m) maxm = v > maxm ? v : maxm;
    by_size.resize(N);
    for (int i = 0; i < N; i++) by_size[i] = i;
    std::stable_sort(by_size.begin(), by_size.end(), [&](int x, int y) { return m[x] > m[y]; });
    return SERBF_OK;
}

WArgs make_args(serbf_engine *e)
{
    const serbf_config &c = e->cfg;
    WArgs a;
    std::memset(&a, 0, sizeof(a));
    a.N = e->N; a.n = e->n; a.ref_mode = c.mode == SERBF_MODE_REFERENCE;
    a.phs = c.phs_order; a.deg = c.poly_deg; a.hv = c.hv_order; a.hvphs = c.hv_phs_order;
    a.eps = c.hv_eps;
    a.x = e->d_x; a.y = e->d_y; a.z = e->d_z;
    a.idx = e->d_idx;
    a.Wx = e->d_Wx; a.Wy = e->d_Wy; a.WL = e->d_WL;
    const bool lap = c.mode == SERBF_MODE_CONFIG && c.poly_deg >= 2;
    a.WL1 = lap ? e->d_WL1 : nullptr;
    a.WL2 = lap ? e->d_WL2 : nullptr;
    a.status = e->d_status;
    a.counter = (int *)e->d_scratch;
    a.nn = e->h_nn.empty() ? nullptr : e->d_nn;
    a.dg = e->h_nn.empty() ? nullptr : e->d_deg;
    a.order = nullptr;
    a.count = e->N;
    return a;
}

int upload_order(serbf_engine *e, const std::vector<int> &order)
{
    if (e->d_worder) (void)hipFree(e->d_worder);
    e->d_worder = nullptr;
    SERBF_HIP(hipMalloc((void **)&e->d_worder, order.size() * sizeof(int)));
    SERBF_HIP(hipMemcpyAsync(e->d_worder, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice,
                             e->stream));
    return SERBF_OK;
}

constexpr int kNClasses = 7;
constexpr int kClasses[kNClasses] = {48, 56, 64, 72, 80, 88, 112};

// Delta^k system size of node i (Gaussian + constant: n + 1; PHS: the gradient size)
int hv_size(const serbf_engine *e, int i, int m_grad)
{
    const serbf_config &c = e->cfg;
    const bool gauss = c.mode == SERBF_MODE_REFERENCE || c.hv_phs_order == 0;
    const int n = e->h_nn.empty() ? e->n : e->h_nn[i];
    return gauss ? n + 1 : m_grad;
}

// One timed run of variant v over the work items `items` (node ids, in the order given).
// LDS kernel: the items are grouped by LDS class (the smallest class that holds the node's
// systems), and each non-empty group is one persistent launch of its class, largest class
// first, so uneven stencils run their small systems at the occupancy of a small class.
// Register-row kernel: the gradient systems grouped by register class (the systems of 97..112
// unknowns by the half-row kernel), then the Delta^k systems likewise, one persistent launch per
// non-empty group.
int run_variant(serbf_engine *e, int v, const std::vector<int> &m, const std::vector<int> &items, float *ms)
{
    const bool hv = e->cfg.hv_order > 0;
    struct Launch {
        int off, cnt, engine, cls, kind;   // engine 0 LDS (cls = LDS class index), 1 rows
    };
    std::vector<Launch> plan;
    std::vector<int> grouped;
    grouped.reserve(2 * items.size());
    std::vector<int> lds_items;
    if (kWVariants[v].rows) {
        for (int kind = 0; kind < (hv ? 2 : 1); kind++)
            for (int ci = kNRowClasses; ci >= 0; ci--) {
                const int off = (int)grouped.size();
                for (int it : items) {
                    const int m1 = m[it], m2 = hv ? hv_size(e, it, m1) : 0;
                    if (reg_class_of(m1) < 0 || reg_class_of(m2) < 0) continue;
                    if (reg_class_of(kind == 0 ? m1 : m2) == ci) grouped.push_back(it);
                }
                if ((int)grouped.size() > off) plan.push_back({off, (int)grouped.size() - off, 1, ci, kind});
            }
        for (int it : items) {
            const int m1 = m[it], m2 = hv ? hv_size(e, it, m1) : 0;
            if (reg_class_of(m1) < 0 || reg_class_of(m2) < 0) lds_items.push_back(it);
        }
    } else {
        lds_items = items;
    }
    for (int ci = kNClasses - 1; ci >= 0; ci--) {
        const int off = (int)grouped.size();
        const int lo = ci > 0 ? kClasses[ci - 1] : 0, hi = kClasses[ci];
        for (int it : lds_items)
            if (m[it] > lo && m[it] <= hi) grouped.push_back(it);
        if ((int)grouped.size() > off) plan.push_back({off, (int)grouped.size() - off, 0, ci, 0});
    }
    int rc = upload_order(e, grouped);
    if (rc) return rc;
    SERBF_HIP(hipEventRecord(e->ev0, e->stream));
    if (!hv) SERBF_HIP(hipMemsetAsync(e->d_WL, 0, (size_t)e->N * e->n * sizeof(double), e->stream));
    for (const Launch &L : plan) {
        WArgs a = make_args(e);
        a.order = e->d_worder + L.off;
        a.count = L.cnt;
        SERBF_HIP(hipMemsetAsync(e->d_scratch, 0, sizeof(int), e->stream));
        if (L.engine == 1) {
            rc = L.cls == kNRowClasses ? launch_rowhalf(e, a, L.kind) : launch_row_class(e, a, L.cls, L.kind);
        } else {
            const int copy = kWVariants[v].copy && variant_fits(kClasses[L.cls], 1) ? 1 : 0;
            rc = launch_class(e, a, kClasses[L.cls], copy);
        }
        if (rc) return rc;
    }
    SERBF_HIP(hipEventRecord(e->ev1, e->stream));
    SERBF_HIP(hipEventSynchronize(e->ev1));
    SERBF_HIP(hipEventElapsedTime(ms, e->ev0, e->ev1));
    return SERBF_OK;
}

}  // namespace

int launch_row_class(serbf_engine *e, const WArgs &a, int ci, int kind)
{
    switch (ci) {
    case 0: return launch_rowlu_48(e, a, kind);
    case 1: return launch_rowlu_56(e, a, kind);
    case 2: return launch_rowlu_64(e, a, kind);
    case 3: return launch_rowlu_72(e, a, kind);
    case 4: return launch_rowlu_80(e, a, kind);
    case 5: return launch_rowlu_88(e, a, kind);
    default: return launch_rowlu_96(e, a, kind);
    }
}

int weights_run(serbf_engine *e)
{
    int maxm = 0;
    std::vector<int> by_size, m;
    int rc = system_sizes(e, maxm, by_size, m);
    if (rc) return rc;
    SERBF_CHECK(maxm <= 112, SERBF_ERR_UNSUPPORTED, "weights: n + polynomial terms must be <= 112");
    rc = alloc_weights(e);
    if (rc) return rc;
    rc = ensure_scratch(e, 64);
    if (rc) return rc;
    int v = e->wvariant;
    if (v < 0 || v >= kNVariants) v = 4;
    std::vector<int> items(e->N);
    for (int i = 0; i < e->N; i++) items[i] = kWVariants[v].sorted ? by_size[i] : i;
    float ms = 0.f;
    if ((rc = run_variant(e, v, m, items, &ms))) return rc;
    e->wstats_ms = ms;
    e->wstats_systems = (double)e->N * (e->cfg.hv_order > 0 ? 2.0 : 1.0);
    e->wstats_variant = v;
    std::vector<int> st(e->N);
    SERBF_HIP(hipMemcpyAsync(st.data(), e->d_status, e->N * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    const serbf_config &c = e->cfg;
    e->have_weights = true;
    e->have_lap = c.mode == SERBF_MODE_CONFIG && c.poly_deg >= 2;
    rk4_free(e);
    e->have_state = false;
    for (int s : st)
        if (s) {
            set_error("weights: at least one singular system (see status)");
            return SERBF_ERR_SINGULAR;
        }
    return SERBF_OK;
}

// Times every variant that fits on the same sample of nodes (every N/sample-th node) and keeps
// the fastest for serbf_weights.  report[3 v] = variant, ms, systems per second (ms < 0: does
// not fit).
int weights_autotune(serbf_engine *e, int sample, double *report, int *nvariants)
{
    int maxm = 0;
    std::vector<int> by_size, m;
    int rc = system_sizes(e, maxm, by_size, m);
    if (rc) return rc;
    SERBF_CHECK(maxm <= 112, SERBF_ERR_UNSUPPORTED, "weights: n + polynomial terms must be <= 112");
    rc = alloc_weights(e);
    if (rc) return rc;
    rc = ensure_scratch(e, 64);
    if (rc) return rc;
    if (sample <= 0 || sample > e->N) sample = e->N;
    // the sample: evenly spaced nodes; its sorted form keeps the same nodes, largest first
    std::vector<int> pick(sample), rank(e->N);
    for (int i = 0; i < sample; i++) pick[i] = (int)((long long)i * e->N / sample);
    for (int i = 0; i < e->N; i++) rank[by_size[i]] = i;
    std::vector<int> sorted_pick(pick);
    std::stable_sort(sorted_pick.begin(), sorted_pick.end(), [&](int x, int y) { return rank[x] < rank[y]; });
    const double systems = (double)sample * (e->cfg.hv_order > 0 ? 2.0 : 1.0);
    int best = 0;
    float best_ms = 1e30f;
    for (int v = 0; v < kNVariants; v++) {
        float ms = -1.f;
        // a copy variant keeps the copy in the classes it fits (it runs without it in the others)
        if (!kWVariants[v].copy || variant_fits(kClasses[0], 1)) {
            // one warm-up run, then the timed one
            rc = run_variant(e, v, m, kWVariants[v].sorted ? sorted_pick : pick, &ms);
            if (!rc) rc = run_variant(e, v, m, kWVariants[v].sorted ? sorted_pick : pick, &ms);
            if (rc) break;
            if (ms < best_ms) { best_ms = ms; best = v; }
        }
        if (report) {
            report[3 * v] = v;
            report[3 * v + 1] = ms;
            report[3 * v + 2] = ms > 0.f ? systems / (ms * 1e-3) : 0.0;
        }
    }
    if (rc) return rc;
    if (nvariants) *nvariants = kNVariants;
    e->wvariant = best;
    return SERBF_OK;
}

// ---- vertical operator (host, tiny) ----
static int host_dgesv(int m, double *A, double *b)
{
    std::vector<int> piv(m);
    for (int k = 0; k < m; k++) {
        int p = k;
        for (int i = k + 1; i < m; i++)
            if (std::fabs(A[i * m + k]) > std::fabs(A[p * m + k])) p = i;
        if (A[p * m + k] == 0.0) return k + 1;
        if (p != k) {
            for (int j = 0; j < m; j++) std::swap(A[k * m + j], A[p * m + j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < m; i++) {
            const double l = A[i * m + k] / A[k * m + k];
            for (int j = k + 1; j < m; j++) A[i * m + j] -= l * A[k * m + j];
            b[i] -= l * b[k];
        }
    }
    for (int i = m - 1; i >= 0; i--) {
        for (int j = i + 1; j < m; j++) b[i] -= A[i * m + j] * b[j];
        b[i] /= A[i * m + i];
    }
    return 0;
}

// Reference mode: the 7-point 6th-order FD1 / dh on levels v-3..v+3, zero outside the
// column (config.h:21-22, rbffd_DMs.c:146-154, 412-427).  Config mode: 1D PHS |z|^kv +
// degree-dv polynomials on the nv levels nearest v (one-sided at the ends), in units of dh.
int vertical_table(const serbf_config &c, int *start, double *W)
{
    const int Nv = c.Nv;
    const double dh = c.htop / Nv;
    if (c.mode == SERBF_MODE_REFERENCE) {
        static const double fd1[7] = {-1.0 / 60.0, 3.0 / 20.0, -3.0 / 4.0, 0.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0};
        for (int v = 0; v < Nv; v++) {
            start[v] = v - 3;
            for (int j = 0; j < 7; j++) W[v * 7 + j] = fd1[j] / dh;
        }
        return 7;
    }
    const int nv = c.nv, kv = c.v_phs_order, dv = c.v_poly_deg, m = nv + dv + 1;
    if (nv < 1 || nv > Nv || nv > kMaxNv || m > 32) return -1;
    for (int v = 0; v < Nv; v++) {
        int s0 = v - nv / 2;
        if (s0 < 0) s0 = 0;
        if (s0 > Nv - nv) s0 = Nv - nv;
        start[v] = s0;
        double zc[kMaxNv];
        for (int j = 0; j < nv; j++) zc[j] = (double)(s0 + j - v);
        std::vector<double> A(m * m, 0.0), b(m, 0.0);
        for (int i = 0; i < nv; i++) {
            for (int j = 0; j < nv; j++) A[i * m + j] = std::pow(std::fabs(zc[i] - zc[j]), kv);
            for (int q = 0; q <= dv; q++) {
                const double p = std::pow(zc[i], q);
                A[i * m + nv + q] = p;
                A[(nv + q) * m + i] = p;
            }
            b[i] = kv * (0.0 - zc[i]) * std::pow(std::fabs(zc[i]), kv - 2);
        }
        for (int q = 0; q <= dv; q++) b[nv + q] = q == 1 ? 1.0 : 0.0;
        if (host_dgesv(m, A.data(), b.data())) return -1;
        for (int j = 0; j < nv; j++) W[v * nv + j] = b[j] / dh;
    }
    return nv;
}

}  // namespace serbf
