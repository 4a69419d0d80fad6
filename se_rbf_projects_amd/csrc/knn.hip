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
#include <mutex>
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

// Grid of cubes of side s over [-1, 1]^3 with a ball of radius s holding ~1.3 n nodes,
// at most gmax cells a side.  The result does not depend on G: every choice gives the exact
// n smallest keys (or the capacity error of knn_query).
static void knn_grid(int N, int n, int gmax, int *G, double *s)
{
    double side = std::sqrt(5.2 * n / (double)N);
    int g = (int)std::floor(2.0 / side);
    if (g < 1) g = 1;
    if (g > gmax) g = gmax;
    *G = g;
    *s = 2.0 / g;
}

// Byte offsets of the search's working arrays in one buffer: count/start/cursor (cells + 1
// ints each), cell_of (N ints), sorted (N double4), the error flag and the scan's temporary.
struct KnnLayout {
    size_t start, cursor, count, cell, sorted, flag, temp, temp_bytes, total;
};

static hipError_t knn_layout(int N, int G, hipStream_t stream, KnnLayout *L)
{
    const size_t cells = (size_t)G * G * G;
    L->temp_bytes = 0;
    const hipError_t rc = hipcub::DeviceScan::ExclusiveSum(nullptr, L->temp_bytes, (int *)nullptr, (int *)nullptr,
                                                           (int)(cells + 1), stream);
    L->start = 0;
    L->cursor = L->start + (cells + 1) * sizeof(int);
    L->count = L->cursor + (cells + 1) * sizeof(int);
    L->cell = L->count + (cells + 1) * sizeof(int);
    L->sorted = (L->cell + (size_t)N * sizeof(int) + 255) & ~(size_t)255;
    L->flag = L->sorted + (size_t)N * sizeof(double4);
    L->temp = (L->flag + 256 + 255) & ~(size_t)255;
    L->total = L->temp + L->temp_bytes + 256;
    return rc;
}

__global__ void knn_zero(int n, int *p)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// The search on `stream` with its working arrays at `base` (laid out by L); waits for the
// stream and returns SERBF_OK or the capacity / size error.
static int knn_search(hipStream_t stream, int N, int n, const double *x, const double *y, const double *z, int G,
                      double s, char *base, const KnnLayout &L, int *idx, double *D)
{
    const size_t cells = (size_t)G * G * G;
    int *start = (int *)(base + L.start), *cursor = (int *)(base + L.cursor), *count = (int *)(base + L.count);
    int *cell_of = (int *)(base + L.cell), *flag = (int *)(base + L.flag);
    double4 *sorted = (double4 *)(base + L.sorted);
    // count[] and flag are cleared by a kernel (not a memset node), so the sequence can be captured
    hipLaunchKernelGGL(knn_zero, dim3(512), dim3(256), 0, stream, (int)(cells + 1), count);
    hipLaunchKernelGGL(knn_zero, dim3(1), dim3(64), 0, stream, 1, flag);
    SERBF_HIP(hipGetLastError());
    const int tb = 256, nb = (N + tb - 1) / tb;
    hipLaunchKernelGGL(knn_count, dim3(nb), dim3(tb), 0, stream, N, x, y, z, G, 1.0 / s, cell_of, count);
    SERBF_HIP(hipGetLastError());
    size_t temp_bytes = L.temp_bytes;
    SERBF_HIP(hipcub::DeviceScan::ExclusiveSum((void *)(base + L.temp), temp_bytes, count, start, (int)(cells + 1),
                                               stream));
    SERBF_HIP(hipMemcpyAsync(cursor, start, (cells + 1) * sizeof(int), hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(knn_scatter, dim3(nb), dim3(tb), 0, stream, N, x, y, z, cell_of, cursor, sorted);
    SERBF_HIP(hipGetLastError());
    const int qb = (N + kKnnWaves - 1) / kKnnWaves;
    hipLaunchKernelGGL(knn_query, dim3(qb), dim3(64 * kKnnWaves), 0, stream, N, n, x, y, z, G, s, 1.0 / s, start,
                       sorted, idx, D, flag);
    SERBF_HIP(hipGetLastError());
    int hflag = 0;
    SERBF_HIP(hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, stream));
    SERBF_HIP(hipStreamSynchronize(stream));
    SERBF_CHECK(hflag == 0, SERBF_ERR_UNSUPPORTED,
                hflag & 1 ? "knn: more candidates than the per-query capacity (clustered nodes)"
                          : "knn: fewer nodes than the stencil size");
    return SERBF_OK;
}

int knn_run(serbf_engine *e, int n)
{
    const int N = e->N;
    int rc = alloc_stencils(e, n);
    if (rc) return rc;
    int G;
    double s;
    knn_grid(N, n, 256, &G, &s);
    KnnLayout L;
    SERBF_HIP(knn_layout(N, G, e->stream, &L));
    rc = ensure_scratch(e, L.total);
    if (rc) return rc;
    rc = knn_search(e->stream, N, n, e->d_x, e->d_y, e->d_z, G, s, (char *)e->d_scratch, L, e->d_idx, e->d_D);
    if (rc) return rc;
    e->have_stencils = true;
    e->have_weights = false;
    return SERBF_OK;
}

}  // namespace serbf

// ---- kNN into caller-held device memory, with no allocation -------------------------------
// The search's working arrays live in one file-scope device array, `scratch`, sized for at
// most kKnnFixedMaxNodes nodes on a grid of at most kKnnFixedMaxGrid cells a side.  Each GPU
// holds its own copy at its own address, found once per device with hipGetSymbolAddress; the
// array's host-side name is only a stand-in and is never handed to a kernel or a copy.
// Every call on one GPU shares that GPU's copy: calls that overlap on one GPU race.
namespace {

constexpr int kKnnFixedMaxNodes = 1 << 20;
constexpr int kKnnFixedMaxGrid = 128;
constexpr size_t kKnnFixedTempBytes = 1 << 20;
constexpr int kKnnFixedMaxDevices = 64;

constexpr size_t knn_fixed_bytes()
{
    const size_t cells = (size_t)kKnnFixedMaxGrid * kKnnFixedMaxGrid * kKnnFixedMaxGrid;
    const size_t cell = 3 * (cells + 1) * sizeof(int);
    const size_t sorted = (cell + (size_t)kKnnFixedMaxNodes * sizeof(int) + 255) & ~(size_t)255;
    const size_t flag = sorted + (size_t)kKnnFixedMaxNodes * sizeof(double4);
    const size_t temp = (flag + 256 + 255) & ~(size_t)255;
    return temp + kKnnFixedTempBytes + 256;
}

constexpr size_t kKnnScratchBytes = knn_fixed_bytes();   // 63,963,904 bytes (61 MiB)

std::mutex g_scratch_mu;
char *g_scratch_dev[kKnnFixedMaxDevices];

}  // namespace

static __device__ __align__(256) unsigned char scratch[kKnnScratchBytes];

// Device address of this GPU's copy of `scratch`; `device` must be current.
static int knn_scratch_base(int device, char **base)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        SERBF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SERBF_CHECK(p != nullptr, SERBF_ERR_HIP, "knn_noalloc: no device address for the working array");
        g_scratch_dev[device] = (char *)p;
    }
    *base = g_scratch_dev[device];
    return SERBF_OK;
}

// True when p is device memory on `device` (a host pointer would fault the kernels).
static bool on_device(const void *p, int device)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return (a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) && a.device == device;
}

extern "C" int serbf_knn_noalloc(int device, int N, const double *x, const double *y, const double *z, int n,
                                 int *idx, double *D)
{
    using namespace serbf;
    SERBF_CHECK(N >= 1 && N <= kKnnFixedMaxNodes, SERBF_ERR_ARG,
                "knn_noalloc: N must be 1..1048576 (the size of the fixed working array)");
    SERBF_CHECK(n >= 1 && n <= N && n <= kMaxStencil, SERBF_ERR_ARG, "knn_noalloc: n out of range");
    SERBF_CHECK(x && y && z && idx && D, SERBF_ERR_ARG, "knn_noalloc: null pointer");
    int ndev = 0;
    SERBF_HIP(hipGetDeviceCount(&ndev));
    SERBF_CHECK(device >= 0 && device < ndev && device < kKnnFixedMaxDevices, SERBF_ERR_ARG,
                "knn_noalloc: no such device");
    SERBF_HIP(hipSetDevice(device));
    SERBF_CHECK(on_device(x, device) && on_device(y, device) && on_device(z, device) && on_device(idx, device) &&
                    on_device(D, device),
                SERBF_ERR_ARG, "knn_noalloc: x, y, z, idx and D must be device memory on the given device");
    char *base = nullptr;
    int rc = knn_scratch_base(device, &base);
    if (rc) return rc;
    int G;
    double s;
    knn_grid(N, n, kKnnFixedMaxGrid, &G, &s);
    KnnLayout L;
    SERBF_HIP(knn_layout(N, G, nullptr, &L));
    SERBF_CHECK(L.total <= kKnnScratchBytes, SERBF_ERR_UNSUPPORTED,
                "knn_noalloc: the search needs more than the fixed working array holds");
    return knn_search(nullptr, N, n, x, y, z, G, s, base, L, idx, D);
}
