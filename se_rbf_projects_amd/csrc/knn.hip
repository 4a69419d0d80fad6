// Bucketed GPU kNN stencil search (gfx950): the body of get_stencils
// (nodesets.c:113-268 of the reference), without its N x N distance slice.
//
// Contract (identical results to the reference on the same inputs):
//   d^2 = fl(fl(fl(dx*dx) + fl(dy*dy)) + fl(dz*dz)), dx = x_i - x_j, every product and
//   sum separately rounded (no FMA contraction, nodesets.c:178-180);
//   the stencil of node i is the n smallest keys (d^2, j) in ascending order, so equal
//   distances keep the lower index first (nodesets.c:228-249, strict '>');
//   D = sqrt(d^2) (nodesets.c:258-267).
//
// Method: nodes are counting-sorted into a G^3 grid of cubes of side s over
// [-1, 1]^3 (s chosen so a ball of radius s holds ~1.3 n nodes).  A wavefront
// handles one query: it scans the (2r+1)^3 cells around the query's cell
// (r = 1, 2, ...), keeps candidates with d^2 <= (r s (1 - 1e-9))^2 in LDS, and
// stops as soon as at least n are kept; every node outside the scanned block is
// farther than r s, so the kept set contains the n smallest keys.  Ranks are
// computed by counting smaller keys, which orders ties by index exactly.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "engine.h"

namespace serbf {

constexpr int kKnnCap = 1024;   // kept candidates per query held in LDS
constexpr int kKnnWaves = 4;    // queries per work-group

__device__ __forceinline__ int cell_coord(double v, double inv_s, int G)
{
    int c = (int)floor((v + 1.0) * inv_s);
    return c < 0 ? 0 : (c >= G ? G - 1 : c);
}

__global__ void knn_count(int N, const double *x, const double *y, const double *z, int G, double inv_s,
                          int *cell_of, int *count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int c = (cell_coord(z[i], inv_s, G) * G + cell_coord(y[i], inv_s, G)) * G + cell_coord(x[i], inv_s, G);
    cell_of[i] = c;
    atomicAdd(&count[c], 1);
}

__global__ void knn_scatter(int N, const double *x, const double *y, const double *z, const int *cell_of,
                            int *cursor, double4 *sorted)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int pos = atomicAdd(&cursor[cell_of[i]], 1);
    sorted[pos] = make_double4(x[i], y[i], z[i], __longlong_as_double((long long)i));
}

// Every product and sum rounded separately: no FMA contraction in this function.
__device__ __forceinline__ double exact_d2(double xi, double yi, double zi, double xj, double yj, double zj)
{
#pragma clang fp contract(off)
    const double dx = xi - xj, dy = yi - yj, dz = zi - zj;
    double d = dx * dx;
    d = d + dy * dy;
    d = d + dz * dz;
    return d;
}

__global__ void __launch_bounds__(64 * kKnnWaves)
knn_query(int N, int n, const double *x, const double *y, const double *z, int G, double s, double inv_s,
          const int *start,
          const double4 *sorted, int *out_idx, double *out_D, int *flag)
{
    __shared__ double kd[kKnnWaves][kKnnCap];
    __shared__ int ki[kKnnWaves][kKnnCap];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = __builtin_amdgcn_readfirstlane(blockIdx.x * kKnnWaves + w);
    if (q >= N) return;
    const double xq = x[q], yq = y[q], zq = z[q];
    const int cx = cell_coord(xq, inv_s, G), cy = cell_coord(yq, inv_s, G), cz = cell_coord(zq, inv_s, G);
    double *D2 = kd[w];
    int *ID = ki[w];
    int count = 0;
    for (int ring = 1;; ring++) {
        const bool all = ring >= G;
        const double rf = ring * s * (1.0 - 1e-9);
        const double rf2 = all ? 1e300 : rf * rf;
        count = 0;
        const int side = 2 * ring + 1;
        const int x0 = cx - ring < 0 ? 0 : cx - ring, x1 = cx + ring >= G ? G - 1 : cx + ring;
        for (int row = 0; row < side * side; row++) {
            const int yy = cy + row % side - ring, zz = cz + row / side - ring;
            if (yy < 0 || yy >= G || zz < 0 || zz >= G) continue;
            const int cb = (zz * G + yy) * G;
            const int b = start[cb + x0], e = start[cb + x1 + 1];
            for (int jb = b; jb < e; jb += 64) {
                const int j = jb + lane;
                double d2 = 0.0;
                int id = -1;
                bool keep = false;
                if (j < e) {
                    const double4 c = sorted[j];
                    d2 = exact_d2(xq, yq, zq, c.x, c.y, c.z);
                    id = (int)__double_as_longlong(c.w);
                    keep = d2 <= rf2;
                }
                const unsigned long long m = __ballot(keep);
                const int off = count + __popcll(m & ((1ull << lane) - 1ull));
                if (keep && off < kKnnCap) {
                    D2[off] = d2;
                    ID[off] = id;
                }
                count += __popcll(m);
            }
        }
        if (count >= n || all || count > kKnnCap) break;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    if (count > kKnnCap || count < n) {
        if (lane == 0) atomicOr(flag, count > kKnnCap ? 1 : 2);
        return;
    }
    for (int c = lane; c < count; c += 64) {
        const double d = D2[c];
        const int id = ID[c];
        int rank = 0;
        for (int o = 0; o < count; o++) {
            const double od = D2[o];
            rank += (od < d) || (od == d && ID[o] < id);
        }
        if (rank < n) {
            out_idx[(size_t)q * n + rank] = id;
            out_D[(size_t)q * n + rank] = sqrt(d);
        }
    }
}

int knn_run(serbf_engine *e, int n)
{
    const int N = e->N;
    int rc = alloc_stencils(e, n);
    if (rc) return rc;
    double s = std::sqrt(5.2 * n / (double)N);
    int G = (int)std::floor(2.0 / s);
    if (G < 1) G = 1;
    if (G > 256) G = 256;
    s = 2.0 / G;
    const size_t cells = (size_t)G * G * G;
    // scratch: count/start (cells+1), cursor (cells+1), cell_of (N), sorted (N double4), flag, scan temp
    size_t temp_bytes = 0;
    SERBF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, (int *)nullptr, (int *)nullptr,
                                               (int)(cells + 1), e->stream));
    const size_t off_start = 0;
    const size_t off_cursor = off_start + (cells + 1) * sizeof(int);
    const size_t off_count = off_cursor + (cells + 1) * sizeof(int);
    const size_t off_cell = off_count + (cells + 1) * sizeof(int);
    const size_t off_sorted = (off_cell + (size_t)N * sizeof(int) + 255) & ~(size_t)255;
    const size_t off_flag = off_sorted + (size_t)N * sizeof(double4);
    const size_t off_temp = (off_flag + 256 + 255) & ~(size_t)255;
    rc = ensure_scratch(e, off_temp + temp_bytes + 256);
    if (rc) return rc;
    char *base = (char *)e->d_scratch;
    int *start = (int *)(base + off_start), *cursor = (int *)(base + off_cursor), *count = (int *)(base + off_count);
    int *cell_of = (int *)(base + off_cell), *flag = (int *)(base + off_flag);
    double4 *sorted = (double4 *)(base + off_sorted);
    SERBF_HIP(hipMemsetAsync(count, 0, (cells + 1) * sizeof(int), e->stream));
    SERBF_HIP(hipMemsetAsync(flag, 0, sizeof(int), e->stream));
    const int tb = 256, nb = (N + tb - 1) / tb;
    hipLaunchKernelGGL(knn_count, dim3(nb), dim3(tb), 0, e->stream, N, e->d_x, e->d_y, e->d_z, G, 1.0 / s, cell_of,
                       count);
    SERBF_HIP(hipGetLastError());
    SERBF_HIP(hipcub::DeviceScan::ExclusiveSum((void *)(base + off_temp), temp_bytes, count, start, (int)(cells + 1),
                                               e->stream));
    SERBF_HIP(hipMemcpyAsync(cursor, start, (cells + 1) * sizeof(int), hipMemcpyDeviceToDevice, e->stream));
    hipLaunchKernelGGL(knn_scatter, dim3(nb), dim3(tb), 0, e->stream, N, e->d_x, e->d_y, e->d_z, cell_of, cursor,
                       sorted);
    SERBF_HIP(hipGetLastError());
    const int qb = (N + kKnnWaves - 1) / kKnnWaves;
    hipLaunchKernelGGL(knn_query, dim3(qb), dim3(64 * kKnnWaves), 0, e->stream, N, n, e->d_x, e->d_y, e->d_z, G, s,
                       1.0 / s, start, sorted, e->d_idx, e->d_D, flag);
    SERBF_HIP(hipGetLastError());
    int hflag = 0;
    SERBF_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    SERBF_CHECK(hflag == 0, SERBF_ERR_UNSUPPORTED,
                hflag & 1 ? "knn: more candidates than the per-query capacity (clustered nodes)"
                          : "knn: fewer nodes than the stencil size");
    e->have_stencils = true;
    e->have_weights = false;
    return SERBF_OK;
}

}  // namespace serbf
