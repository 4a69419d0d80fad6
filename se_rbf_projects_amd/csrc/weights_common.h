// Shared device pieces of the batched weight solves (weights.hip, rowlu*.hip): the kernel
// arguments, the matrix entries and right-hand sides of the per-node saddle systems
// (rbffd_DMs.c:156-261 of the reference, and the config-mode systems), and the node's
// rotated, scaled stencil frame.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

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
__device__ inline double lap_gauss(double r, double eps, int k)
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

// right-hand-side columns per system (gradient system: d/dx', d/dy', Delta, Delta^2)
constexpr int kRhs = 4;

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

// Right-hand sides of row i of the current system (kind 0: the gradient system's d/dx', d/dy',
// Delta, Delta^2; kind 1: the Delta^k system's one column), zero past what the system uses
__device__ __forceinline__ void rhs_of(int i, int n, int kind, const WArgs &a, const double *X, int MAXN, double b[kRhs])
{
    b[0] = b[1] = b[2] = b[3] = 0.0;
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
}

// Gather the node's first n stencil neighbours, rotate them into the node's frame
// (rbffd_DMs.c:159-168) and, in config mode, shift to the centre and scale by the stencil
// radius; X[3][MAXN] in LDS.  Returns the scale (1 in reference mode).
__device__ inline double stencil_frame(const WArgs &a, int node, int n, double *X, int MAXN, int lane)
{
    double H[9];
    householder_dev(a.x[node], a.y[node], a.z[node], !a.ref_mode, H);
    const int *ip = a.idx + (size_t)node * a.n;
    for (int s = lane; s < n; s += 64) {
        const int q = ip[s];
        const double px = a.x[q], py = a.y[q], pz = a.z[q];
        for (int r = 0; r < 3; r++) X[r * MAXN + s] = H[3 * r] * px + H[3 * r + 1] * py + H[3 * r + 2] * pz;
    }
    wave_sync();
    double scale = 1.0;
    if (!a.ref_mode) {
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
    return scale;
}

// One launch of the register-row kernel of class ci (rowlu.h) for system kind (0 gradient,
// 1 Delta^k) over a.count work items.
int launch_row_class(serbf_engine *e, const WArgs &a, int ci, int kind);

}  // namespace serbf
