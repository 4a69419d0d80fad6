// Batched per-node RBF-FD weight solves (gfx950), the body of get_rbffd_DMs
// (rbffd_DMs.c:98-282 of the reference).
//
// One persistent launch factors every node's saddle system.  A wavefront owns one
// system at a time: it gathers the stencil, rotates it with the node's
// Householder reflector, assembles [Phi P; P^T 0] and the right-hand sides in LDS
// (lane = row, rows lane + 64 r), factors it with partial pivoting (first maximal
// |a_ik|, full-row swaps as LAPACK dgesv), solves, and applies one step of
// iterative refinement whose residual b - A x re-evaluates A from the
// coordinates (the matrix itself is never stored in HBM).  Work is pulled from a
// global counter so uneven system sizes balance across the chip.
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

#include "device_util.h"
#include "engine.h"

namespace serbf {

struct WArgs {
    int N, n, ref_mode;
    int phs, deg, hv, hvphs;
    double eps;
    const double *x, *y, *z;
    const int *idx;
    double *Wx, *Wy, *WL;
    double *WL1, *WL2;   // config mode: Delta and Delta^2 weights from the gradient system (may be null)
    int *status;
    int *counter;
    const int *nn, *dg;  // per-node stencil size and polynomial degree (null: n, deg for every node)
    const int *order;    // work item -> node (null: identity)
    int count;           // work items
};

// x^k for a small integer k >= 0 by repeated squaring (the PHS orders and monomial degrees)
__device__ __forceinline__ double ipow(double x, int k)
{
    double r = 1.0, b = x;
    while (k > 0) {
        if (k & 1) r *= b;
        b *= b;
        k >>= 1;
    }
    return r;
}
__device__ __forceinline__ double phs_pow(double r, int k) { return k >= 0 ? ipow(r, k) : pow(r, (double)k); }

// 2D Delta^k of the Gaussian exp(-(eps r)^2) by the Laguerre recursion (rbffd_DMs.c:314-328)
__device__ double lap_gauss(double r, double eps, int k)
{
    double epr2 = pow(eps * r, 2.0);
    double pm = 1.0, p = 4 * epr2 - 4;
    if (k == 0) p = 1.0;
    for (int i = 1; i < k; i++) {
        double pn = 4 * (epr2 - (2 * i + 1)) * p - (16 * pow((double)i, 2.0)) * pm;
        pm = p;
        p = pn;
    }
    return pow(eps, 2.0 * k) * p * exp(-pow(r * eps, 2.0));
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// LDS layout of one wavefront's system (doubles):
//   A   [MAXM][LD]  augmented matrix [A | B] (LD >= MAXM + 4, odd)
//   X   [3][MAXN]   stencil coordinates in the solve frame
//   B0  [MAXM][4]   right-hand sides (kept for the residual)
//   XS  [MAXM][4]   solution
//   piv [MAXM]      pivot rows (int)
// right-hand-side columns per system (gradient system: d/dx', d/dy', Delta, Delta^2)
constexpr int kRhs = 4;

//   A0  [MAXM][MAXM] (COPY variant) the assembled matrix, for the refinement residual
template <int MAXM, bool COPY>
struct Lds {
    static constexpr int LD = ((MAXM + 4) | 1);
    static constexpr int MAXN = MAXM;
    static constexpr int doubles = MAXM * LD + 3 * MAXN + 2 * MAXM * kRhs + (COPY ? MAXM * MAXM : 0);
    static constexpr int bytes = doubles * 8 + MAXM * 4;
};

// Matrix entry A_ij of the current system (rows/cols < m); kind 0: gradient PHS system,
// kind 1: Delta^k system, Gaussian + constant (reference mode, or config mode with
// hv_phs_order 0; a.eps is the shape parameter in the solve frame) or PHS r^K + monomials.
__device__ __forceinline__ double entry(int i, int j, int n, int m, int kind, const WArgs &a, const double *X,
                                        int MAXN)
{
    const bool gauss = kind == 1 && (a.ref_mode || a.hvphs == 0);
    if (i < n && j < n) {
        const double dx = X[j] - X[i], dy = X[MAXN + j] - X[MAXN + i], dz = X[2 * MAXN + j] - X[2 * MAXN + i];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (gauss) return exp(-pow(r * a.eps, 2.0));
        return phs_pow(r, kind == 0 ? a.phs : a.hvphs);
    }
    if (i >= n && j >= n) return 0.0;
    const int node = i < n ? i : j;
    const int t = i < n ? j - n : i - n;
    if (gauss) return 1.0;
    // monomial t in the order x^(p-q) y^q, p = 0..deg, q = 0..p
    int p = 0;
    while ((p + 1) * (p + 2) / 2 <= t) p++;
    const int q = t - p * (p + 1) / 2;
    return ipow(X[node], p - q) * ipow(X[MAXN + node], q);
}

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
        double b[kRhs] = {0.0, 0.0, 0.0, 0.0};
        if (i < n) {
            // derivative at the centre (stencil node 0) of the basis function centred at node i
            const double dx = X[0] - X[i], dy = X[MAXN] - X[MAXN + i], dz = X[2 * MAXN] - X[2 * MAXN + i];
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (kind == 0) {
                // d/dx', d/dy', and (config mode) Delta, Delta^2 of r^k at the centre:
                // Delta^j r^k = prod_{i<j} (k - 2i)^2 r^(k - 2j) (2D radial Laplacian)
                const int k = a.phs;
                const double K = (double)k;
                b[0] = k * dx * phs_pow(r, k - 2);
                b[1] = k * dy * phs_pow(r, k - 2);
                b[2] = K * K * phs_pow(r, k - 2);
                b[3] = K * K * (K - 2.0) * (K - 2.0) * phs_pow(r, k - 4);
            } else if (a.ref_mode || a.hvphs == 0) {
                b[0] = lap_gauss(r, a.eps, a.hv);
            } else {
                const double K = (double)a.hvphs;
                double c = 1.0;
                for (int jj = 0; jj < a.hv; jj++) c *= (K - 2 * jj) * (K - 2 * jj);
                b[0] = c * phs_pow(r, a.hvphs - 2 * a.hv);
            }
        } else if (kind == 0) {
            // d/dx' of x' and d/dy' of y' at the origin of the tangent plane; Delta and Delta^2 of
            // the monomial x^(p-q) y^q there (2 for x^2, y^2; 24, 8, 24 for x^4, x^2 y^2, y^4).
            // Delta^k (k >= 3) of a monomial of degree < 2k vanishes there.
            const int t = i - n;
            int p = 0;
            while ((p + 1) * (p + 2) / 2 <= t) p++;
            const int q = t - p * (p + 1) / 2;
            b[0] = (t == 1) ? 1.0 : 0.0;
            b[1] = (t == 2) ? 1.0 : 0.0;
            b[2] = (p == 2 && (q == 0 || q == 2)) ? 2.0 : 0.0;
            b[3] = p == 4 ? ((q == 0 || q == 4) ? 24.0 : (q == 2 ? 8.0 : 0.0)) : 0.0;
        }
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
        double H[9];
        householder_dev(a.x[node], a.y[node], a.z[node], !a.ref_mode, H);
        // gather and rotate the stencil (rbffd_DMs.c:159-168)
        const int *ip = a.idx + (size_t)node * stride;
        for (int s = lane; s < n; s += 64) {
            const int q = ip[s];
            const double px = a.x[q], py = a.y[q], pz = a.z[q];
            for (int r = 0; r < 3; r++) X[r * MAXN + s] = H[3 * r] * px + H[3 * r + 1] * py + H[3 * r + 2] * pz;
        }
        wave_sync();
        double scale = 1.0;
        if (!a.ref_mode) {
            // shift to the centre, scale by the stencil radius
            const double c0 = X[0], c1 = X[MAXN], c2 = X[2 * MAXN];
            wave_sync();
            double rmax = 0.0;
            for (int s = lane; s < n; s += 64) {
                const double dx = X[s] - c0, dy = X[MAXN + s] - c1, dz = X[2 * MAXN + s] - c2;
                X[s] = dx; X[MAXN + s] = dy; X[2 * MAXN + s] = dz;
                rmax = fmax(rmax, sqrt(dx * dx + dy * dy + dz * dz));
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) rmax = fmax(rmax, __shfl_xor(rmax, off, 64));
            scale = rmax;
            wave_sync();
            for (int s = lane; s < n; s += 64) {
                X[s] /= scale; X[MAXN + s] /= scale; X[2 * MAXN + s] /= scale;
            }
            wave_sync();
        }
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
    int copy;      // 1: matrix copy in LDS
    int sorted;    // 1: largest systems first
};
constexpr WVariant kWVariants[4] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};

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

// system size of every node, the largest one and the order of the nodes by decreasing size
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

// One timed run of variant v over the work items `items` (node ids, in the order given).  The
// items are grouped by LDS class (the smallest class that holds the node's system), and each
// non-empty group is one persistent launch of its class, largest class first: uneven stencils
// then run their small systems at the occupancy of a small class.
int run_variant(serbf_engine *e, int v, const std::vector<int> &m, const std::vector<int> &items, float *ms)
{
    std::vector<int> grouped;
    grouped.reserve(items.size());
    int start[kNClasses] = {};
    for (int ci = kNClasses - 1; ci >= 0; ci--) {
        start[ci] = (int)grouped.size();
        const int lo = ci > 0 ? kClasses[ci - 1] : 0, hi = kClasses[ci];
        for (int it : items)
            if (m[it] > lo && m[it] <= hi) grouped.push_back(it);
    }
    int rc = upload_order(e, grouped);
    if (rc) return rc;
    SERBF_HIP(hipEventRecord(e->ev0, e->stream));
    for (int ci = kNClasses - 1; ci >= 0; ci--) {
        const int off = start[ci], cnt = (ci > 0 ? start[ci - 1] : (int)grouped.size()) - off;
        if (cnt <= 0) continue;
        WArgs a = make_args(e);
        a.order = e->d_worder + off;
        a.count = cnt;
        const int copy = kWVariants[v].copy && variant_fits(kClasses[ci], 1) ? 1 : 0;
        SERBF_HIP(hipMemsetAsync(e->d_scratch, 0, sizeof(int), e->stream));
        if ((rc = launch_class(e, a, kClasses[ci], copy))) return rc;
    }
    SERBF_HIP(hipEventRecord(e->ev1, e->stream));
    SERBF_HIP(hipEventSynchronize(e->ev1));
    SERBF_HIP(hipEventElapsedTime(ms, e->ev0, e->ev1));
    return SERBF_OK;
}

}  // namespace

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
    if (v < 0 || v > 3) v = 0;
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
    for (int v = 0; v < 4; v++) {
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
    if (nvariants) *nvariants = 4;
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
