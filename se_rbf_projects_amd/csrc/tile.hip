// LDS-tiled RK4 stage kernel (gfx950): the right-hand side of eval_partSVDotGradSV
// (rbffd_DMs.c:331-443 of the reference) with the RK4 axpys fused, as in rk4.hip,
// but with the neighbour gathers staged through LDS.
//
// A work-group owns a tile of 16 consecutive nodes (nodes are in space-filling
// order, so a tile is spatially compact).  At set-up, tile_plan_kernel sorts the
// tile's 16 x n stencil indices in LDS and stores their union (the distinct
// neighbour columns, ~10 per node instead of n) and, per stencil entry, a 16-bit
// position in that union.  Per stage the work-group walks the state in 256-byte
// chunks of every column (two tracers x 32 levels in FP32, two tracers x 16 levels
// in FP64, whole 128-byte lines of each tracer column): it copies the chunk of every
// union column and the own columns' chunk +- 8 levels into LDS once, then every
// node's stencil sum reads its neighbours from LDS.  Each global column chunk is
// thus fetched once per tile instead of once per stencil entry.
//
// Lanes: 4 nodes per wavefront, 16 lanes per node, 16 bytes per lane (4 FP32 or
// 2 FP64 levels of one tracer).
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "engine.h"
#include "stage_common.h"

namespace serbf {

constexpr int kTileNodes = 16;
constexpr int kTileKeys = 2048;   // 16 nodes x n <= 128

// ---------------------------------------------------------------- set-up
__global__ void __launch_bounds__(256) tile_plan_kernel(int N, int n, int umax, const int *idx, int *uni, int *ucount,
                                                        unsigned short *lidx, int *flag)
{
    __shared__ int keys[kTileKeys];
    __shared__ int uniq[kTileKeys];
    __shared__ int cnt;
    const int t = blockIdx.x, tid = threadIdx.x;
    const int p0 = t * kTileNodes;
    const int m = kTileNodes * n;
    int M = 1;
    while (M < m) M <<= 1;
    for (int k = tid; k < M; k += blockDim.x) {
        int v = INT_MAX;
        if (k < m) {
            const int p = p0 + k / n;
            if (p < N) v = idx[(size_t)p * n + k % n];
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
    for (int i = tid; i < c && i < umax; i += blockDim.x) uni[(size_t)t * umax + i] = uniq[i];
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
        lidx[(size_t)p * n + k % n] = (unsigned short)lo;
    }
}

// ---------------------------------------------------------------- stage kernel
template <typename T, int NTR, int CHB>
struct TileGeom {
    static constexpr int S = 64;
    static constexpr int BPL = CHB / 16;                            // bytes per lane
    static constexpr int VPL = BPL / sizeof(T);                     // values per lane
    static constexpr int TRC = (NTR == 1 || CHB == 128) ? 1 : 2;    // tracers per chunk
    static constexpr int CH = CHB / TRC / sizeof(T);                // levels per chunk
    static constexpr int LPT = 16 / TRC;                            // lanes per tracer
    static constexpr int LCH = S / CH;                              // level chunks
    static constexpr int NCH = (NTR / TRC) * LCH;                   // chunks per column
    static constexpr int HALO = CH + 16;                            // own-column window
    static_assert(LPT * VPL == CH, "lane geometry");
    // dynamic LDS: union chunks, union ids, metadata (a, b, l, local index), halo windows
    // (umax is a multiple of 4, npad of 4: every array below stays 16-byte aligned)
    __host__ __device__ static size_t bytes(int umax, int npad, int xy)
    {
        return (size_t)umax * CHB + (size_t)umax * 4 + (size_t)kTileNodes * npad * ((2 + xy) * sizeof(T) + 2) +
               (size_t)kTileNodes * TRC * HALO * sizeof(T) + 64;
    }
};

struct TileArgs {
    const int *uni;              // union ids, ustride per tile
    const int *ucount;
    const unsigned short *lidx;  // position of each stencil entry in its tile's union
    int ustride;                 // allocation stride of uni
    int umax;                    // largest union present (LDS capacity)
    int npad, ntiles;
};

template <typename T, int BPL = 16>
__device__ __forceinline__ Vec4<T> load_vpl(const T *p)
{
    Vec4<T> r;
    if constexpr (BPL == 8) {
        if constexpr (sizeof(T) == 4) {
            const float2 x = *(const float2 *)p;
            r.v[0] = x.x; r.v[1] = x.y;
        } else {
            r.v[0] = *p;
        }
    } else if constexpr (sizeof(T) == 4) {
        const float4 x = *(const float4 *)p;
        r.v[0] = x.x; r.v[1] = x.y; r.v[2] = x.z; r.v[3] = x.w;
    } else {
        const double2 x = *(const double2 *)p;
        r.v[0] = x.x; r.v[1] = x.y; r.v[2] = T(0); r.v[3] = T(0);
    }
    return r;
}

template <typename T, int BPL = 16>
__device__ __forceinline__ void store_vpl(T *p, const Vec4<T> &r)
{
    if constexpr (BPL == 8) {
        if constexpr (sizeof(T) == 4) *(float2 *)p = make_float2(r.v[0], r.v[1]);
        else *p = r.v[0];
    } else if constexpr (sizeof(T) == 4) {
        *(float4 *)p = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
    } else {
        *(double2 *)p = make_double2(r.v[0], r.v[1]);
    }
}

// Largest tile union the kernel stages (columns); larger unions use the
// column-vector stage kernel.
constexpr int kTileUcap = 256;

template <typename T, int NTR, int STAGE, int HV, int CHB>
__global__ void __launch_bounds__(256) tile_kernel(StageArgs<T> a, TileArgs ta, int xy)
{
    using G = TileGeom<T, NTR, CHB>;
    constexpr int S = G::S, VPL = G::VPL, TRC = G::TRC, CH = G::CH, LPT = G::LPT, LCH = G::LCH, HALO = G::HALO;
    constexpr int BPL = G::BPL;
    constexpr int ULD = kTileUcap * 16 / 256;                     // union parts per thread
    constexpr int HV4 = HALO / (16 / sizeof(T));                  // 16-byte vectors per halo window
    constexpr int HLD = (kTileNodes * TRC * HV4 + 255) / 256;     // halo vectors per thread
    using V16 = float4;
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ T vtab[S * kMaxNv];
    __shared__ int vst[S];
    __shared__ T slotv[kTileNodes][2];
    const int tid = threadIdx.x;
    const int tile = xcd_remap(blockIdx.x, a.nblk);
    const int p0 = tile * kTileNodes;
    const int npad = ta.npad, n = a.n, umax = ta.umax;
    const int u = ta.ucount[tile];

    unsigned char *base = lds;
    unsigned char *ucol = base;
    base += (size_t)umax * CHB;
    int *uid = (int *)base;
    base += (size_t)umax * 4;
    T *ma = (T *)base;
    T *mb = xy ? ma + kTileNodes * npad : ma;
    T *ml = mb + kTileNodes * npad;
    base += (size_t)kTileNodes * npad * (2 + xy) * sizeof(T);
    T *halo = (T *)base;
    base += (size_t)kTileNodes * TRC * HALO * sizeof(T);
    unsigned short *mi = (unsigned short *)base;

    const T *ts = a.tscal + ((size_t)(*a.base + a.step_off) * 4 + (STAGE == 0 ? 0 : STAGE - 1)) * kTimeScalars;

    // ---- per-block tables: vertical stencils, union ids, node velocities ----
    for (int i = tid; i < S * kMaxNv; i += 256) vtab[i] = a.vW[i];
    if (tid < S) vst[tid] = a.vstart[tid];
    for (int i = tid; i < u; i += 256) uid[i] = ta.uni[(size_t)tile * ta.ustride + i];
    if (tid < kTileNodes && !xy) {
        int p = p0 + tid;
        if (p >= a.N) p = a.N - 1;
        T ux, uy;
        node_velocity<T>(a, a.basis + (size_t)p * kBasis, ts, ux, uy);
        slotv[tid][0] = ux;
        slotv[tid][1] = uy;
    }
    __syncthreads();

    // ---- stencil metadata: all loads issued before use ----
    {
        constexpr int MPT = kTileNodes * 64 / 256;   // entries per thread for npad <= 64
        T rx[MPT], ry[MPT], rl[MPT];
        unsigned short ri[MPT];
#pragma unroll
        for (int r = 0; r < MPT; r++) {
            const int e = tid + 256 * r;
            rx[r] = ry[r] = rl[r] = T(0);
            ri[r] = 0;
            if (e < kTileNodes * npad) {
                const int sl = e / npad, s = e - sl * npad;
                int p = p0 + sl;
                if (p >= a.N) p = a.N - 1;
                const size_t o = (size_t)p * n + (s < n ? s : 0);
                rx[r] = a.Wx[o];
                ry[r] = a.Wy[o];
                if (HV) rl[r] = a.WL[o];
                ri[r] = ta.lidx[o];
            }
        }
#pragma unroll
        for (int r = 0; r < MPT; r++) {
            const int e = tid + 256 * r;
            if (e < kTileNodes * npad) {
                const int sl = e / npad, s = e - sl * npad;
                const bool live = s < n;
                if (xy) {
                    ma[e] = live ? rx[r] : T(0);
                    mb[e] = live ? ry[r] : T(0);
                } else {
                    ma[e] = live ? slotv[sl][0] * rx[r] + slotv[sl][1] * ry[r] : T(0);
                }
                ml[e] = live ? rl[r] : T(0);
                mi[e] = ri[r];
            }
        }
    }

    const int lane = tid & 63, w = tid >> 6;
    const int slot = 4 * w + (lane >> 4), j = lane & 15;
    const int trl = j / LPT, lv = (j % LPT) * VPL;
    int p = p0 + slot;
    const bool node_ok = p < a.N;
    if (!node_ok) p = a.N - 1;
    const T *B = a.basis + (size_t)p * kBasis;

    // register staging of one chunk: union parts and own-column halo vectors
    V16 ru[ULD];
    V16 rh[HLD];
    auto fetch = [&](int c) {
        const int tb = c / LCH, lb = c % LCH;
        const int tr0 = tb * TRC, l0 = lb * CH;
#pragma unroll
        for (int r = 0; r < ULD; r++) {
            const int k = tid + 256 * r;
            if (k < u * 16) {
                const int col = k >> 4, part = k & 15;
                const int t2 = part / LPT, l2 = (part % LPT) * VPL;
                const T *src = a.qin + ((size_t)uid[col] * NTR + tr0 + t2) * S + l0 + l2;
                if constexpr (BPL == 16) ru[r] = *(const float4 *)src;
                else {
                    const float2 x = *(const float2 *)src;
                    ru[r] = make_float4(x.x, x.y, 0.f, 0.f);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < HLD; r++) {
            const int k = tid + 256 * r;
            rh[r] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (k < kTileNodes * TRC * HV4) {
                const int i = k / (TRC * HV4), rr = k - i * (TRC * HV4);
                const int t2 = rr / HV4, lev = l0 - 8 + (rr - t2 * HV4) * (16 / (int)sizeof(T));
                int pi = p0 + i;
                if (pi >= a.N) pi = a.N - 1;
                if (lev >= 0 && lev < S) rh[r] = *(const float4 *)(a.qin + ((size_t)pi * NTR + tr0 + t2) * S + lev);
            }
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int r = 0; r < ULD; r++) {
            const int k = tid + 256 * r;
            if (k < u * 16) {
                if constexpr (BPL == 16) *(float4 *)(ucol + (size_t)k * 16) = ru[r];
                else *(float2 *)(ucol + (size_t)k * 8) = make_float2(ru[r].x, ru[r].y);
            }
        }
#pragma unroll
        for (int r = 0; r < HLD; r++) {
            const int k = tid + 256 * r;
            if (k < kTileNodes * TRC * HV4) *(float4 *)((unsigned char *)halo + (size_t)k * 16) = rh[r];
        }
    };

    fetch(0);
    commit();
    for (int c = 0; c < G::NCH; c++) {
        const int tb = c / LCH, lb = c % LCH;
        const int tr0 = tb * TRC, l0 = lb * CH;
        __syncthreads();                 // chunk c (and the metadata) visible
        if (c + 1 < G::NCH) fetch(c + 1);  // in flight while chunk c is computed

        // horizontal stencil from LDS
        T acc_a[VPL], acc_b[VPL], acc_l[VPL];
#pragma unroll
        for (int k = 0; k < VPL; k++) acc_a[k] = acc_b[k] = acc_l[k] = T(0);
        const T *mra = ma + slot * npad, *mrb = mb + slot * npad, *mrl = ml + slot * npad;
        const unsigned short *mri = mi + slot * npad;
        const unsigned char *urow = ucol + j * BPL;
        if (xy) {
            for (int s0 = 0; s0 < npad; s0 += 4) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int s = s0 + q;
                    const Vec4<T> x = load_vpl<T, BPL>((const T *)(urow + (size_t)mri[s] * CHB));
                    const T wa = mra[s], wb = mrb[s], wl = mrl[s];
#pragma unroll
                    for (int k = 0; k < VPL; k++) {
                        acc_a[k] = fma(wa, x.v[k], acc_a[k]);
                        acc_b[k] = fma(wb, x.v[k], acc_b[k]);
                        if (HV) acc_l[k] = fma(wl, x.v[k], acc_l[k]);
                    }
                }
            }
        } else {
            for (int s0 = 0; s0 < npad; s0 += 4) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int s = s0 + q;
                    const Vec4<T> x = load_vpl<T, BPL>((const T *)(urow + (size_t)mri[s] * CHB));
                    const T wa = mra[s], wl = mrl[s];
#pragma unroll
                    for (int k = 0; k < VPL; k++) {
                        acc_a[k] = fma(wa, x.v[k], acc_a[k]);
                        if (HV) acc_l[k] = fma(wl, x.v[k], acc_l[k]);
                    }
                }
            }
        }

        // vertical stencil from the halo window, velocity, fused RK4 update
        const T *hw = halo + (slot * TRC + trl) * HALO;
        const size_t my = ((size_t)p * NTR + tr0 + trl) * S + l0 + lv;
        T f[VPL], fh[VPL], self[VPL];
#pragma unroll
        for (int k = 0; k < VPL; k++) {
            const int v = l0 + lv + k;
            self[k] = hw[8 + lv + k];
            T z = T(0);
            if (v < a.Nv) {
                const int st = vst[v];
                const T *wv = vtab + v * kMaxNv;
                for (int q = 0; q < a.nv; q++) z = fma(wv[q], hw[st + q - l0 + 8], z);
            }
            T upx, upy, upz;
            level_velocity<T, S>(a, B, ts, p, v, upx, upy, upz);
            const T r_inv = a.lev[v];
            const T adv = xy ? upx * (acc_a[k] * r_inv) + upy * (acc_b[k] * r_inv) : acc_a[k] * r_inv;
            T fv = a.alpha * (adv + upz * z), hv = T(0);
            if (HV) {
                const T cf = a.hv_mode ? a.gamma[my + k] : a.hv_coef;
                if (HV == 1) fv += cf * acc_l[k];
                else hv = T(6) * cf * acc_l[k];
            }
            const bool live = v < a.Nv;
            f[k] = live ? fv : T(0);
            fh[k] = live ? hv : T(0);
        }
        if (node_ok) {
            if (STAGE == 0) {
                for (int k = 0; k < VPL; k++) a.F[my + k] = f[k];
            } else if (STAGE == 1) {
                Vec4<T> F1, o;
#pragma unroll
                for (int k = 0; k < VPL; k++) {
                    F1.v[k] = f[k] + fh[k];
                    o.v[k] = l0 + lv + k < a.Nv ? self[k] + a.hdt * f[k] : T(0);
                }
                store_vpl<T, BPL>(a.F + my, F1);
                store_vpl<T, BPL>(a.qout + my, o);
            } else if (STAGE == 2 || STAGE == 3) {
                const Vec4<T> Fo = load_vpl<T, BPL>(a.F + my), qt = load_vpl<T, BPL>(a.qT + my);
                Vec4<T> Fn, o;
                const T h = STAGE == 2 ? a.hdt : a.dt;
#pragma unroll
                for (int k = 0; k < VPL; k++) {
                    Fn.v[k] = Fo.v[k] + T(2) * f[k];
                    o.v[k] = l0 + lv + k < a.Nv ? qt.v[k] + h * f[k] : T(0);
                }
                store_vpl<T, BPL>(a.F + my, Fn);
                store_vpl<T, BPL>(a.qout + my, o);
            } else {
                const Vec4<T> Fo = load_vpl<T, BPL>(a.F + my), qt = load_vpl<T, BPL>(a.qT + my);
                Vec4<T> o;
#pragma unroll
                for (int k = 0; k < VPL; k++)
                    o.v[k] = l0 + lv + k < a.Nv ? qt.v[k] + a.dt6 * (Fo.v[k] + T(1) * f[k]) : T(0);
                store_vpl<T, BPL>(a.qout + my, o);
            }
        }
        if (c + 1 < G::NCH) {
            __syncthreads();             // every wavefront finished reading chunk c
            commit();
        }
    }
}

// ---------------------------------------------------------------- host
int tile_plan(serbf_engine *e)
{
    if (e->have_plan) return SERBF_OK;
    SERBF_CHECK(e->n <= kTileKeys / kTileNodes, SERBF_ERR_UNSUPPORTED, "tile plan: n > 128");
    const int ntiles = (e->N + kTileNodes - 1) / kTileNodes;
    const int umax = kTileNodes * e->n < 1024 ? kTileNodes * e->n : 1024;
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
    e->have_plan = true;
    // tiles whose union does not fit in LDS use the column-vector stage kernel
    e->plan_failed = flag != 0 || e->tile_umax > kTileUcap || ((e->n + 3) & ~3) > 64 ||
                     TileGeom<double, 4, 256>::bytes(e->tile_umax, (e->n + 3) & ~3, 1) > 150 * 1024;
    return SERBF_OK;
}

bool tile_supported(const serbf_engine *e)
{
    return e->S == 64 && (e->ntr == 1 || e->ntr == 4) && e->n <= 128 && !(e->have_plan && e->plan_failed);
}

template <typename T, int NTR, int CHB>
static void launch_tile_t(serbf_engine *e, int stage, int hv, bool xy, StageArgs<T> a)
{
    using G = TileGeom<T, NTR, CHB>;
    TileArgs ta;
    ta.uni = e->d_tuni;
    ta.ucount = e->d_tcount;
    ta.lidx = e->d_lidx;
    ta.ustride = e->tile_ustride;
    ta.umax = e->tile_umax;
    ta.npad = (e->n + 3) & ~3;
    ta.ntiles = e->ntiles;
    a.nblk = e->ntiles;
    const int x = xy ? 1 : 0;
    const size_t lds = G::bytes(ta.umax, ta.npad, x);
    const dim3 grid(e->ntiles), block(256);
#define SERBF_T(ST, HVV)                                                                                    \
    do {                                                                                                  \
        (void)hipFuncSetAttribute((const void *)tile_kernel<T, NTR, ST, HVV, CHB>,                             \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                  \
        hipLaunchKernelGGL((tile_kernel<T, NTR, ST, HVV, CHB>), grid, block, lds, e->stream, a, ta, x);        \
    } while (0)
    switch (stage * 3 + hv) {
    case 0: SERBF_T(0, 0); break;
    case 1: SERBF_T(0, 1); break;
    case 3: SERBF_T(1, 0); break;
    case 4: SERBF_T(1, 1); break;
    case 5: SERBF_T(1, 2); break;
    case 6: SERBF_T(2, 0); break;
    case 7: SERBF_T(2, 1); break;
    case 9: SERBF_T(3, 0); break;
    case 10: SERBF_T(3, 1); break;
    case 12: SERBF_T(4, 0); break;
    default: SERBF_T(4, 1); break;
    }
#undef SERBF_T
}

template <typename T>
void launch_tile(serbf_engine *e, int stage, int hv, bool xy, const StageArgs<T> &a)
{
    if (e->tile_chunk == 128) {
        if (e->ntr == 1) launch_tile_t<T, 1, 128>(e, stage, hv, xy, a);
        else launch_tile_t<T, 4, 128>(e, stage, hv, xy, a);
    } else {
        if (e->ntr == 1) launch_tile_t<T, 1, 256>(e, stage, hv, xy, a);
        else launch_tile_t<T, 4, 256>(e, stage, hv, xy, a);
    }
}

template void launch_tile<float>(serbf_engine *, int, int, bool, const StageArgs<float> &);
template void launch_tile<double>(serbf_engine *, int, int, bool, const StageArgs<double> &);

}  // namespace serbf
