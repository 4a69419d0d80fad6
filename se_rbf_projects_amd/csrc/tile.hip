// Tile plan of the matrix-core stage kernel (mtile.hip), built once per stencil set.
//
// A tile is 16 consecutive nodes (nodes are in space-filling order, so a tile is
// spatially compact).  tile_plan_kernel sorts the tile's 16 x n stencil indices together
// with the tile's own 16 node ids in LDS and stores their union (the distinct columns,
// about 126 at n = 55 instead of 880) and, per stencil entry, its 16-bit position in that
// union.  The tile's own nodes come first, in node order (union positions 0..15), the
// other columns follow in ascending id order: the vector stage kernel takes the tile's own
// columns from its first step of 16 union columns.  The stage kernels read each union
// column once per tile and scatter the stencil weights into a dense 16 x union block.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "engine.h"
#include "stage_common.h"

namespace serbf {

constexpr int kTileNodes = 16;
constexpr int kTileKeys = 4096;   // 16 nodes x (n + 1) with n <= 128

// ---------------------------------------------------------------- set-up
__global__ void __launch_bounds__(256) tile_plan_kernel(int N, int n, int umax, const int *idx, int *uni, int *ucount,
                                                        unsigned short *lidx, int *flag)
{
    __shared__ int keys[kTileKeys];
    __shared__ int uniq[kTileKeys];
    __shared__ int cnt;
    const int t = blockIdx.x, tid = threadIdx.x;
    const int p0 = t * kTileNodes;
    const int own = N - p0 < kTileNodes ? N - p0 : kTileNodes;   // the tile's own nodes
    const int m = kTileNodes * n;
    int M = 1;
    while (M < m + kTileNodes) M <<= 1;
    for (int k = tid; k < M; k += blockDim.x) {
        int v = INT_MAX;
        if (k < m) {
            const int p = p0 + k / n;
            if (p < N) v = idx[(size_t)p * n + k % n];
        } else if (k < m + own) {
            v = p0 + (k - m);   // own ids, also where a stencil leaves its centre out
        }
        keys[k] = v;
    }
    __syncthreads();
    for (int size = 2; size <= M; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = tid; k < M; k += blockDim.x) {
                const int j = k ^ stride;
                if (j > k) {
                    const bool up = (k & size) == 0;
                    const int a = keys[k], b = keys[j];
                    if ((a > b) == up) {
                        keys[k] = b;
                        keys[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    if (tid == 0) {
        int c = 0;
        for (int k = 0; k < M; k++) {
            const int v = keys[k];
            if (v == INT_MAX) break;
            if (k == 0 || v != keys[k - 1]) uniq[c++] = v;
        }
        cnt = c;
        ucount[t] = c;
        if (c > umax) atomicOr(flag, 1);
    }
    __syncthreads();
    const int c = cnt;
    // the own ids p0 .. p0 + own - 1 are consecutive, so they form one run [o, o + own) of the
    // sorted union; they move to the front and the columns below them shift up by `own`
    int o = 0, hi = c;
    while (o < hi) {
        const int mid = (o + hi) >> 1;
        if (uniq[mid] < p0) o = mid + 1;
        else hi = mid;
    }
    auto front = [&](int s) { return s < o ? s + own : (s < o + own ? s - o : s); };
    for (int i = tid; i < c; i += blockDim.x) {
        const int j = front(i);
        if (j < umax) uni[(size_t)t * umax + j] = uniq[i];
    }
    for (int k = tid; k < m; k += blockDim.x) {
        const int p = p0 + k / n;
        if (p >= N) continue;
        const int v = idx[(size_t)p * n + k % n];
        int lo = 0, hi = c - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (uniq[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        lidx[(size_t)p * n + k % n] = (unsigned short)front(lo);
    }
}

// ---------------------------------------------------------------- host
int tile_plan(serbf_engine *e)
{
    if (e->have_plan) return SERBF_OK;
    SERBF_CHECK(e->n <= kTileKeys / kTileNodes - 1 && e->n <= kMaxStencil, SERBF_ERR_UNSUPPORTED,
                "tile plan: n > 128");
    const int ntiles = (e->N + kTileNodes - 1) / kTileNodes;
    const int umax = kTileNodes * (e->n + 1) < 1024 ? kTileNodes * (e->n + 1) : 1024;
    for (void **p : {(void **)&e->d_tuni, (void **)&e->d_tcount, (void **)&e->d_lidx})
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    SERBF_HIP(hipMalloc((void **)&e->d_tuni, (size_t)ntiles * umax * sizeof(int)));
    SERBF_HIP(hipMalloc((void **)&e->d_tcount, (size_t)ntiles * sizeof(int)));
    SERBF_HIP(hipMalloc((void **)&e->d_lidx, (size_t)e->N * e->n * sizeof(unsigned short)));
    int rc = ensure_scratch(e, 64);
    if (rc) return rc;
    SERBF_HIP(hipMemsetAsync(e->d_scratch, 0, sizeof(int), e->stream));
    hipLaunchKernelGGL(tile_plan_kernel, dim3(ntiles), dim3(256), 0, e->stream, e->N, e->n, umax, e->d_idx, e->d_tuni,
                       e->d_tcount, e->d_lidx, (int *)e->d_scratch);
    SERBF_HIP(hipGetLastError());
    std::vector<int> cnt(ntiles);
    int flag = 0;
    SERBF_HIP(hipMemcpyAsync(cnt.data(), e->d_tcount, ntiles * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipMemcpyAsync(&flag, e->d_scratch, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    int mx = 0;
    for (int c : cnt) mx = c > mx ? c : mx;
    e->tile_ustride = umax;
    e->tile_umax = (mx + 3) & ~3;
    e->ntiles = ntiles;
    // a union wider than the plan's stride (flag) is recorded as too wide for the tiles,
    // which then leave the stage to the column-vector kernel (mtile_supported)
    if (flag) e->tile_umax = INT_MAX & ~3;
    e->have_plan = true;
    return SERBF_OK;
}

}  // namespace serbf
