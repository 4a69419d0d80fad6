// Stage kernels on the matrix cores (gfx950): the horizontal stencil of a 16-node tile as a
// dense product, W (16 nodes x U union columns) x Q (U x levels), on v_mfma_f32_16x16x4f32
// (FP32) or v_mfma_f64_16x16x4f64 (FP64).
//
// A tile is 16 consecutive (space-filling-curve ordered) nodes.  The tile plan (tile.hip) gives
// the union of their stencils (U <= kMTileUcap = 320 columns; a stencil set with a wider tile
// runs on the column kernel) and, for every stencil entry, its position in the union.  Per stage a work-group scatters the combined advection weights u_x' Dx + u_y' Dy
// of its 16 nodes (and the Delta^k weights when the stage applies hyperviscosity) into dense
// 16 x U blocks in LDS, multiplies them by the union's columns on the matrix cores, and applies
// the vertical stencil, the velocity and the fused RK4 update to the result.
// Two kernels read the union columns straight from memory as the B operand, each column once
// per tile: mtile2_kernel (stage-kernel version 5: FP64, and FP32 shapes other than 4 tracers
// of 64 levels) and mvec_kernel (version 7, the FP32 default for 4 tracers of 64 levels, below).
// MFMA operand layout (16x16x4, lane l, c = l % 16, g = l / 16):
//   A (16 x 4): A[c][g]     B (4 x 16): B[g][c]
//   D (16 x 16): FP32 D[4g + i][c], FP64 D[g + 4i][c], i = 0..3
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "engine.h"
#include "stage_common.h"
#include "tile_common.h"

namespace serbf {

namespace {

constexpr int kMTileNodes = 16;
// largest tile union (columns) the matrix-core kernel takes: 320 covers the n = 84 stencils of
// uneven node sets (266 columns at most on jittered level-5 nodes); the LDS then stays within
// 84 KB (FP64)
constexpr int kMTileUcap = 320;

template <typename T>
struct Acc4;
template <>
struct Acc4<float> {
    typedef v4f type;
    // tile row (node) of result value i of a lane in row group g
    __device__ static int row(int g, int i) { return 4 * g + i; }
};
template <>
struct Acc4<double> {
    typedef v4d type;
    __device__ static int row(int g, int i) { return g + 4 * i; }
};

// ---------------------------------------------------------------------------
// Version 5 (FP32 on v_mfma_f32_16x16x4f32, FP64 on v_mfma_f64_16x16x4f64): the same
// 16-node tiles and dense weight blocks, with the state read straight from global memory as
// the MFMA B operand.
//
// Lane (c, g) of wavefront w loads level 16w + c of union column k0 + g: the four wavefronts
// together read each union column once per tile (the stencil-by-stencil gather of the column
// kernel reads it once per stencil entry, about 5.5 times as often).  Nothing of the state is
// staged in LDS, so the main loop has no barrier; LDS holds only the dense weight blocks and the
// union offsets (about 34 KB at 256 union columns), which leaves room for four work-groups per
// CU.  With TPW tracers per wavefront, each step of 16 union columns issues 4 x TPW independent
// loads (4 column groups x TPW tracers) and 8 x TPW MFMAs into 2 x TPW independent accumulators
// (TPW tracers x advection and Delta^k): 8 loads and 16 MFMAs at TPW = 2.  After the loop the
// LDS is reused for the tile's own columns (zero-padded by 8 levels on either side) for the
// vertical stencil and the RK4 update.
constexpr int kOwnLD = 80;   // own-column stride: 8 + 64 + 8 levels

template <typename T>
__host__ __device__ inline int mtile2_lds_bytes(int up, bool hv, int ntr)
{
    const int wld = up + 1;
    const int wreg = 16 * wld * (hv ? 2 : 1);
    const int oreg = kMTileNodes * ntr * kOwnLD;
    return (32 + (wreg > oreg ? wreg : oreg)) * (int)sizeof(T) + up * (int)sizeof(int);
}

// NTR tracers (4 or 1) of 64 levels per node; TPW tracers per wavefront: NTR / TPW groups of
// 4 wavefronts, wavefront w taking level block w % 4 of tracers TPW (w / 4) .. TPW (w / 4) + TPW - 1
// Occupancy: 8 wavefronts per SIMD in FP32 (64 registers, 4 work-groups per CU) and 4 in FP64
// (128 registers with a 20-byte spill, 2 work-groups per CU).  At 655362 x 60 x 4 this took the
// FP32 stage from 1.43 to 1.36 ms and the FP64 stage from 2.96 ms (3 wavefronts per SIMD, one
// work-group per CU) to 2.41 ms.
template <typename T, int STAGE, int HV, int NTR, int TPW>
__global__ void __launch_bounds__(256 * (NTR / TPW)) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 8 : 4)))
mtile2_kernel(StageArgs<T> a, MTilePlan tp, int up)
{
    extern __shared__ __align__(16) unsigned char m2_lds[];
    typedef typename Acc4<T>::type VT;
    constexpr int S = 64, NTH = 256 * (NTR / TPW);
    const int wld = up + 1;
    const int wreg = 16 * wld * (HV ? 2 : 1), oreg = kMTileNodes * NTR * kOwnLD;
    T *sv = (T *)m2_lds;
    T *W = sv + 32;
    T *WLt = W + 16 * wld;
    T *own = W;                                       // reused after the main loop
    int *uoff = (int *)(W + (wreg > oreg ? wreg : oreg));
    const int tid = threadIdx.x, lane = tid & 63, w = (tid >> 6) & 3, t0 = (tid >> 8) * TPW;
    const int c = lane & 15, g = lane >> 4;
    const int tile = xcd_remap(blockIdx.x, a.nblk);
    const int p0 = tile * kMTileNodes;
    const int U = tp.ucount[tile];
    const int u16 = (U + 15) & ~15;
    const int n = a.n;
    const T *ts = a.tscal + ((size_t)(*a.base + a.step_off) * 4 + (STAGE == 0 ? 0 : STAGE - 1)) * kTimeScalars;
    const int *uni = tp.uni + (size_t)tile * tp.ustride;

    // ---- node velocities, union offsets, dense weight blocks ----
    // FP64: the tile's weight rows (contiguous, rows p0 .. p0 + 15) are loaded into registers
    // before the first barrier, with the velocities and union offsets, so the set-up waits for
    // one memory round trip instead of two.  At 4 wavefronts per SIMD that latency shows: the
    // FP64 stage at 655362 x 60 x 4 went from 2.42 to 2.35 ms.  At 8 wavefronts per SIMD the
    // FP32 stage hides it and gained nothing (1.365 -> 1.374 ms), so FP32 loads after the barrier.
    constexpr bool kPre = sizeof(T) == 8;
    constexpr int kWE = kPre ? (kMTileNodes * kMaxStencil + NTH - 1) / NTH : 1;
    T wxr[kWE], wyr[kWE], wlr[kWE];
    int kr[kWE];
    if (kPre) {
        const int ne = (p0 + kMTileNodes <= a.N ? kMTileNodes : a.N - p0) * n;
#pragma unroll
        for (int r = 0; r < kWE; r++) {
            const int e = tid + r * NTH;
            kr[r] = -1;
            if (e < ne) {
                const size_t o = (size_t)p0 * n + e;
                kr[r] = tp.lidx[o];
                wxr[r] = a.Wx[o];
                wyr[r] = a.Wy[o];
                if (HV) wlr[r] = a.WL[o];
            }
        }
    }
    if (tid < kMTileNodes) {
        const int p = p0 + tid < a.N ? p0 + tid : a.N - 1;
        T ux, uy;
        node_velocity<T>(a, a.basis + (size_t)p * kBasis, ts, ux, uy);
        sv[2 * tid] = ux;
        sv[2 * tid + 1] = uy;
    }
    for (int k = tid; k < u16; k += NTH) uoff[k] = (k < U ? uni[k] : uni[0]) * (NTR * S);
    for (int e = tid; e < wreg; e += NTH) W[e] = T(0);
    __syncthreads();
    if (kPre) {
#pragma unroll
        for (int r = 0; r < kWE; r++)
            if (kr[r] >= 0) {
                const int t = (tid + r * NTH) / n;
                W[t * wld + kr[r]] = sv[2 * t] * wxr[r] + sv[2 * t + 1] * wyr[r];
                if (HV) WLt[t * wld + kr[r]] = wlr[r];
            }
    } else {
        for (int e = tid; e < kMTileNodes * n; e += NTH) {
            const int t = e / n, s = e - t * n;
            const int p = p0 + t;
            if (p < a.N) {
                const size_t o = (size_t)p * n + s;
                const int k = tp.lidx[o];
                W[t * wld + k] = sv[2 * t] * a.Wx[o] + sv[2 * t + 1] * a.Wy[o];
                if (HV) WLt[t * wld + k] = a.WL[o];
            }
        }
    }
    __syncthreads();

    // ---- D (16 nodes x 16 levels of wavefront w) = W (16 x u16) x Q (u16 x 16), per tracer ----
    VT acc[TPW], accl[TPW];
#pragma unroll
    for (int t = 0; t < TPW; t++) {
        acc[t] = VT{T(0), T(0), T(0), T(0)};
        accl[t] = VT{T(0), T(0), T(0), T(0)};
    }
    const T *qw = a.qin + t0 * S + 16 * w + c;
    const T *wa = W + c * wld + g;
    const T *wl = WLt + c * wld + g;
    for (int k0 = 0; k0 < u16; k0 += 16) {
        int off[4];
        T b[4][TPW], av[4], al[4];
#pragma unroll
        for (int j = 0; j < 4; j++) off[j] = uoff[k0 + 4 * j + g];
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int t = 0; t < TPW; t++) b[j][t] = qw[off[j] + t * S];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            av[j] = wa[k0 + 4 * j];
            if (HV) al[j] = wl[k0 + 4 * j];
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int t = 0; t < TPW; t++) {
                acc[t] = mfma4(av[j], b[j][t], acc[t]);
                if (HV) accl[t] = mfma4(al[j], b[j][t], accl[t]);
            }
    }

    // ---- own columns of the tile in LDS, 8 zero levels on either side ----
    __syncthreads();   // every wavefront is done with W and uoff
    for (int e = tid; e < kMTileNodes * NTR * (S / 4); e += NTH) {
        const int col = e >> 4, l4 = (e & 15) * 4;
        const int t = col / NTR, tr = col - t * NTR;
        const int p = p0 + t < a.N ? p0 + t : a.N - 1;
        const Vec4<T> x = load4<T>(a.qin + ((size_t)p * NTR + tr) * S + l4);
        T *d = own + col * kOwnLD + 8 + l4;
        d[0] = x.v[0]; d[1] = x.v[1]; d[2] = x.v[2]; d[3] = x.v[3];
    }
    for (int e = tid; e < kMTileNodes * NTR * 16; e += NTH) {
        const int col = e >> 4, k = e & 15;
        own[col * kOwnLD + (k < 8 ? k : S + k)] = T(0);
    }
    __syncthreads();

    // ---- vertical stencil, velocity, RK4 update: 4 nodes x 4 tracers x 1 level per lane ----
    const int v = 16 * w + c;
    const bool live = v < a.Nv;
    const T r_inv = a.lev[v];
    const int vst = a.vstart[live ? v : 0];
    T z[4][TPW];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int t = 0; t < TPW; t++) z[i][t] = T(0);
    if (live)
        for (int j = 0; j < a.nv; j++) {
            const T wv = a.vW[v * kMaxNv + j];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int t = 0; t < TPW; t++)
                    z[i][t] = fma(wv, own[(Acc4<T>::row(g, i) * NTR + t0 + t) * kOwnLD + 8 + vst + j], z[i][t]);
        }
    // every load of the update is issued before its first store: the stores may alias the
    // loaded arrays as far as the compiler knows, so loads placed after a store would wait for
    // one memory round trip per output
    T Fv[4][TPW], Qv[4][TPW], Gv[4][TPW], upz[4];
    size_t mys[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int pr = p0 + Acc4<T>::row(g, i);
        const int p = pr < a.N ? pr : a.N - 1;
        mys[i] = ((size_t)p * NTR + t0) * S + v;
        T upx, upy;
        level_velocity<T, S>(a, a.basis + (size_t)p * kBasis, ts, p, v, upx, upy, upz[i]);
        (void)upx;
        (void)upy;
    }
    // the coefficient source is the same for the whole launch
    if (HV && a.hv_mode == 2) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int t = 0; t < TPW; t++) Gv[i][t] = a.hv_coef * (T)(float)a.s16[mys[i] + t * S];
    } else if (HV && a.hv_mode) {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int t = 0; t < TPW; t++) Gv[i][t] = a.gamma[mys[i] + t * S];
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int t = 0; t < TPW; t++) Gv[i][t] = a.hv_coef;
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int t = 0; t < TPW; t++) {
            Fv[i][t] = STAGE >= 2 ? a.F[mys[i] + t * S] : T(0);
            Qv[i][t] = STAGE >= 2 ? a.qT[mys[i] + t * S] : T(0);
        }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int nd = Acc4<T>::row(g, i);
        const int p = p0 + nd;
        if (p >= a.N) continue;
#pragma unroll
        for (int t = 0; t < TPW; t++) {
            const size_t my = ((size_t)p * NTR + t0 + t) * S + v;
            T f = a.alpha * (acc[t][i] * r_inv + upz[i] * z[i][t]);
            if (HV) f += Gv[i][t] * accl[t][i];
            if (!live) f = T(0);
            if (STAGE == 0) {
                a.F[my] = f;
            } else if (STAGE == 1) {
                a.F[my] = f;
                a.qout[my] = live ? own[(nd * NTR + t0 + t) * kOwnLD + 8 + v] + a.hdt * f : T(0);
            } else if (STAGE == 2 || STAGE == 3) {
                const T h = STAGE == 2 ? a.hdt : a.dt;
                a.F[my] = Fv[i][t] + T(2) * f;
                a.qout[my] = live ? Qv[i][t] + h * f : T(0);
            } else {
                a.qout[my] = live ? Qv[i][t] + a.dt6 * (Fv[i][t] + T(1) * f) : T(0);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Version 7 (FP32, 4 tracers): the same tiles and dense weight blocks, with every state access
// a 16-byte vector.  Wavefront w of a work-group takes tracer w, all 64 levels; lane (c, g) loads
// levels 4c .. 4c + 3 of union column k0 + 4j + g (one float4) and uses its value j' as the B
// operand of level block j': block j' computes levels 4n + j' (n = 0..15) of the 16 nodes, so one
// vector load feeds 4 level blocks x 2 operators = 8 MFMAs and each step of 16 union columns is
// 4 loads and 32 MFMAs into 8 independent accumulators.  The dense blocks put node t in MFMA row
// 4 (t % 4) + t / 4, so lane (c, g) ends with nodes g + 4i (i = 0..3) at levels 4c .. 4c + 3 --
// the same four nodes whose columns it loaded in the first step, where the tile plan lists the
// tile's own nodes first.  Those own columns give the vertical stencil without another load:
// the levels of the lanes c - 1 and c + 1 come by DPP row shifts (zero beyond the row, which are
// the zero ghosts below level 0 and above level 63), and a per-level table of 12 coefficients
// over levels 4c - 4 .. 4c + 7 (built in LDS at set-up from the vertical stencil) covers the
// one-sided rows at the ends.  The update's loads and stores are float4 per node.
__host__ __device__ inline int mvec_lds_bytes(int up, bool hv)
{
    const int wld = up + 1;
    return (32 + 16 * wld * (hv ? 2 : 1) + kMaxLevels * kVecWin) * 4 + up * 4;
}

// 4 wavefronts per SIMD (128 registers), 4 work-groups of 4 wavefronts per CU.  HVM is the
// hyperviscosity coefficient's source: 0 none, 1 the constant hv_coef, 2 hv_coef times the
// per-point FP16 multiplier (the closure), 3 the per-point FP32 coefficient; one source per
// instantiation keeps the update's registers to what that source needs.
template <int STAGE, int HVM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
mvec_kernel(StageArgs<float> a, MTilePlan tp, int up)
{
    constexpr bool HV = HVM != 0;
    extern __shared__ __align__(16) unsigned char mv_lds[];
    constexpr int S = 64, NTR = 4, NTH = 256;
    const int wld = up + 1;
    float *sv = (float *)mv_lds;                      // [16][2] node velocities
    float *W = sv + 32;                               // [16][wld] advection weights (row vec_row(t))
    float *WLt = W + 16 * wld;                        // [16][wld] Delta^k weights
    float *Ct = WLt + (HV ? 16 * wld : 0);            // [64][kVecWin] vertical coefficients
    int *uoff = (int *)(Ct + kMaxLevels * kVecWin);   // [up] element offsets of the union columns
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int tile = xcd_remap(blockIdx.x, a.nblk);
    const int p0 = tile * kMTileNodes;
    const int U = tp.ucount[tile];
    const int u16 = (U + 15) & ~15;
    const int n = a.n;
    const float *ts = a.tscal + ((size_t)(*a.base + a.step_off) * 4 + (STAGE == 0 ? 0 : STAGE - 1)) * kTimeScalars;
    const int *uni = tp.uni + (size_t)tile * tp.ustride;

    // the tile's own columns (union positions 0..15) do not need the union list: their loads
    // go out first and land during the set-up
    const unsigned lo4 = (unsigned)(w * S + 4 * c) * 4u;   // byte offset of the lane's levels in a node
    // In the last tile of a node count that is not a multiple of 16 the positions past its own
    // nodes hold other union columns: those lanes load the column the first step needs (their
    // nodes are past N, so their vertical stencil is never used).
    float4 own[4];
    if (p0 + kMTileNodes <= a.N) {
#pragma unroll
        for (int j = 0; j < 4; j++) own[j] = ld_at<float4>(a.qin, (unsigned)(p0 + 4 * j + g) * (NTR * S * 4u) + lo4);
    } else {
        const int nown = a.N - p0, U0 = tp.ucount[tile];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = 4 * j + g;
            const int id = k < nown ? p0 + k : tp.uni[(size_t)tile * tp.ustride + (k < U0 ? k : 0)];
            own[j] = ld_at<float4>(a.qin, (unsigned)id * (NTR * S * 4u) + lo4);
        }
    }

    // ---- set-up: weight rows into registers, velocities, union offsets, vertical table ----
    constexpr int kWE = (kMTileNodes * kMaxStencil + NTH - 1) / NTH;
    float wxr[kWE], wyr[kWE], wlr[kWE];
    int kr[kWE];
    const int ne = (p0 + kMTileNodes <= a.N ? kMTileNodes : a.N - p0) * n;
#pragma unroll
    for (int r = 0; r < kWE; r++) {
        const int e = tid + r * NTH;
        kr[r] = -1;
        if (e < ne) {
            const size_t o = (size_t)p0 * n + e;
            kr[r] = tp.lidx[o];
            wxr[r] = a.Wx[o];
            wyr[r] = a.Wy[o];
            if (HV) wlr[r] = a.WL[o];
        }
    }
    if (tid < kMTileNodes) {
        const int p = p0 + tid < a.N ? p0 + tid : a.N - 1;
        float ux, uy;
        node_velocity<float>(a, a.basis + (size_t)p * kBasis, ts, ux, uy);
        sv[2 * tid] = ux;
        sv[2 * tid + 1] = uy;
    }
    for (int k = tid; k < u16; k += NTH) uoff[k] = (k < U ? uni[k] : uni[0]) * (NTR * S * 4);   // bytes
    build_vertical_table(Ct, a.Nv, a.nv, a.vstart, a.vW, tid, NTH);
    for (int e = tid; e < 16 * wld * (HV ? 2 : 1); e += NTH) W[e] = 0.f;
    __syncthreads();
    const float rn = 1.f / (float)n;
#pragma unroll
    for (int r = 0; r < kWE; r++)
        if (kr[r] >= 0) {
            // (e + 1/2) / n is at least 1/(2n) away from an integer, far above the rounding
            const int t = (int)(((float)(tid + r * NTH) + 0.5f) * rn);
            const int row = vec_row(t);
            W[row * wld + kr[r]] = sv[2 * t] * wxr[r] + sv[2 * t + 1] * wyr[r];
            if (HV) WLt[row * wld + kr[r]] = wlr[r];
        }
    __syncthreads();

    // ---- D (16 nodes x 64 levels of tracer w) = W (16 x u16) x Q (u16 x 64) ----
    const float *wa = W + c * wld + g;
    const float *wl = WLt + c * wld + g;
    v4f acc[4], accl[4];
    auto step = [&](int k0, const float4 (&b)[4]) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float av = wa[k0 + 4 * j];
            const float al = HV ? wl[k0 + 4 * j] : 0.f;
            const float bv[4] = {b[j].x, b[j].y, b[j].z, b[j].w};
#pragma unroll
            for (int lb = 0; lb < 4; lb++) {
                acc[lb] = mfma4(av, bv[lb], acc[lb]);
                if (HV) accl[lb] = mfma4(al, bv[lb], accl[lb]);
            }
        }
    };
    // The B loads run one step ahead of the MFMAs: the four loads of step k0 + 16 are issued
    // before the 32 MFMAs of step k0, and a scheduling barrier keeps the compiler from sinking
    // them (left to itself it issues two loads, waits for them, and so on: two memory round
    // trips per step).
    auto load_step = [&](int k0, float4 (&b)[4]) {
#pragma unroll
        for (int j = 0; j < 4; j++) b[j] = ld_at<float4>(a.qin, (unsigned)uoff[k0 + 4 * j + g] + lo4);
    };
    // first step: the tile's own columns (union positions 0..15, node 4j + g in load j)
    float4 bn[4];
    if (u16 > 16) load_step(16, bn);
    // The vertical term starts the advection accumulator, f = alpha r_inv (A q + (u_z / r_inv)
    // Dz q) + c L q: the stencil of levels 4c .. 4c + 3 of node g + 4i over the window of levels
    // 4c - 4 .. 4c + 7, the own column and its neighbour lanes' levels.
    const int v0 = 4 * c;
    {
        const float4 rinv = *(const float4 *)(a.lev + v0);
        const float4 wprof = *(const float4 *)(a.lev + S + v0);
        const float wt = a.flow == VEL_DEFORM ? ts[0] : 0.f;
        const float sc[4] = {wt * wprof.x / rinv.x, wt * wprof.y / rinv.y, wt * wprof.z / rinv.z,
                             wt * wprof.w / rinv.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float z[4];
            vertical4(Ct, v0, own[i], z);
#pragma unroll
            for (int lb = 0; lb < 4; lb++) acc[lb][i] = sc[lb] * z[lb];
        }
#pragma unroll
        for (int lb = 0; lb < 4; lb++) accl[lb] = v4f{0.f, 0.f, 0.f, 0.f};
    }
    // the update's loads (F, q_T, coefficient)
    float4 Fv[4], Qv[4], Gv[HVM == 3 ? 4 : 1];
    ushort4 Hv[HVM == 2 ? 4 : 1];
    // element index of the lane's first level of node g + 4i (4-byte values; FP16 multipliers at half)
    auto myi = [&](int i) -> unsigned {
        const int pr = p0 + g + 4 * i;
        const int p = pr < a.N ? pr : a.N - 1;
        return (unsigned)p * (NTR * S) + (unsigned)(w * S + v0);
    };
    auto load_update = [&]() {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const unsigned m = myi(i);
            if constexpr (HVM == 2) Hv[i] = ld_nt_h4(a.s16, m * 2u);
            if constexpr (HVM == 3) Gv[i] = ld_at<float4>(a.gamma, m * 4u);
            Fv[i] = STAGE >= 2 ? ld_nt4(a.F, m * 4u) : make_float4(0.f, 0.f, 0.f, 0.f);
            Qv[i] = STAGE >= 1 ? ld_nt4(a.qT, m * 4u) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    __builtin_amdgcn_sched_barrier(0);
    step(0, own);
    // two steps per iteration with alternating buffers (no register copies)
    for (int k0 = 16; k0 < u16; k0 += 32) {
        float4 bm[4];
        if (k0 + 16 < u16) load_step(k0 + 16, bm);
        __builtin_amdgcn_sched_barrier(0);
        step(k0, bn);
        if (k0 + 16 >= u16) break;
        if (k0 + 32 < u16) load_step(k0 + 32, bn);
        __builtin_amdgcn_sched_barrier(0);
        step(k0 + 16, bm);
    }
    load_update();

    // ---- coefficient and RK4 update: nodes g + 4i, levels 4c .. 4c + 3 ----
    const float4 rinv2 = *(const float4 *)(a.lev + v0);
    const float rv[4] = {rinv2.x, rinv2.y, rinv2.z, rinv2.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (p0 + g + 4 * i >= a.N) continue;
        float gv[4] = {a.hv_coef, a.hv_coef, a.hv_coef, a.hv_coef};
        if constexpr (HVM == 2) {
            const unsigned short hh[4] = {Hv[i].x, Hv[i].y, Hv[i].z, Hv[i].w};
#pragma unroll
            for (int lb = 0; lb < 4; lb++) gv[lb] = a.hv_coef * (float)__builtin_bit_cast(_Float16, hh[lb]);
        } else if constexpr (HVM == 3) {
            gv[0] = Gv[i].x; gv[1] = Gv[i].y; gv[2] = Gv[i].z; gv[3] = Gv[i].w;
        }
        const float fv[4] = {Fv[i].x, Fv[i].y, Fv[i].z, Fv[i].w};
        const float qv[4] = {Qv[i].x, Qv[i].y, Qv[i].z, Qv[i].w};
        float fo[4], qo[4];
#pragma unroll
        for (int lb = 0; lb < 4; lb++) {
            float f = a.alpha * (acc[lb][i] * rv[lb]);
            if (HV) f += gv[lb] * accl[lb][i];
            const bool live = v0 + lb < a.Nv;
            if (!live) f = 0.f;
            if (STAGE <= 1) {
                fo[lb] = f;
                qo[lb] = live ? qv[lb] + a.hdt * f : 0.f;
            } else if (STAGE == 2 || STAGE == 3) {
                const float h = STAGE == 2 ? a.hdt : a.dt;
                fo[lb] = fv[lb] + 2.f * f;
                qo[lb] = live ? qv[lb] + h * f : 0.f;
            } else {
                fo[lb] = 0.f;
                qo[lb] = live ? qv[lb] + a.dt6 * (fv[lb] + 1.f * f) : 0.f;
            }
        }
        const unsigned m = myi(i) * 4u;
        if (STAGE <= 3) st_nt4(a.F, m, make_float4(fo[0], fo[1], fo[2], fo[3]));
        if (STAGE >= 1) st_nt4(a.qout, m, make_float4(qo[0], qo[1], qo[2], qo[3]));
    }
}

template <int STAGE, int HVM>
void launch_vec_s(serbf_engine *e, const StageArgs<float> &a, const MTilePlan &tp, int up, int bytes)
{
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)mvec_kernel<STAGE, HVM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr = true;
    }
    hipLaunchKernelGGL((mvec_kernel<STAGE, HVM>), dim3(e->ntiles), dim3(256), bytes, e->stream, a, tp, up);
}

template <int STAGE>
void launch_vec_m(serbf_engine *e, int hvm, const StageArgs<float> &a, const MTilePlan &tp, int up, int bytes)
{
    switch (hvm) {
    case 0: launch_vec_s<STAGE, 0>(e, a, tp, up, bytes); break;
    case 1: launch_vec_s<STAGE, 1>(e, a, tp, up, bytes); break;
    case 2: launch_vec_s<STAGE, 2>(e, a, tp, up, bytes); break;
    default: launch_vec_s<STAGE, 3>(e, a, tp, up, bytes); break;
    }
}

int launch_mvec(serbf_engine *e, int stage, int hv, const StageArgs<float> &a, const MTilePlan &tp)
{
    const int up = (e->tile_umax + 15) & ~15;
    const int bytes = mvec_lds_bytes(up, hv != 0);
    SERBF_CHECK(bytes <= 160 * 1024, SERBF_ERR_UNSUPPORTED, "vector stage kernel: tile union too large");
    const int hvm = !hv ? 0 : (a.hv_mode == 2 ? 2 : (a.hv_mode ? 3 : 1));
    switch (stage) {
    case 0: launch_vec_m<0>(e, hvm, a, tp, up, bytes); break;
    case 1: launch_vec_m<1>(e, hvm, a, tp, up, bytes); break;
    case 2: launch_vec_m<2>(e, hvm, a, tp, up, bytes); break;
    case 3: launch_vec_m<3>(e, hvm, a, tp, up, bytes); break;
    default: launch_vec_m<4>(e, hvm, a, tp, up, bytes); break;
    }
    SERBF_HIP(hipGetLastError());
    return SERBF_OK;
}

// The vector kernel needs 4 tracers (a wavefront each) and a vertical stencil whose taps for
// level l lie within levels 4 (l / 4) - 4 .. 4 (l / 4) + 7 (n_v <= 7 centred, one-sided at the ends).
bool mvec_supported(const serbf_engine *e)
{
    // 4 tracers of 64 levels, state arrays below 4 GiB (32-bit byte offsets)
    return e->ntr == 4 && (size_t)e->N * 4 * 64 * 4 < (1ull << 32) && vertical_window_fits(e);
}

// tracers per wavefront of the version-5 kernel: 2 (8 wavefronts per work-group).  Measured at
// 655362 x 60 x 4 with the batched update loads, before the occupancy caps: FP32 1.44 ms per
// stage (4 tracers per wavefront: 1.44 ms), FP64 3.08 ms (4 tracers: 3.57 ms, one wavefront
// per SIMD for lack of registers)
template <typename T>
constexpr int kTPW = 2;

template <typename T, int STAGE, int HV, int NTR, int TPW>
void launch_two_n(serbf_engine *e, const StageArgs<T> &a, const MTilePlan &tp, int up, int bytes)
{
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)mtile2_kernel<T, STAGE, HV, NTR, TPW>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    hipLaunchKernelGGL((mtile2_kernel<T, STAGE, HV, NTR, TPW>), dim3(e->ntiles), dim3(256 * (NTR / TPW)), bytes,
                       e->stream, a, tp, up);
}

template <typename T, int STAGE, int HV>
void launch_two(serbf_engine *e, const StageArgs<T> &a, const MTilePlan &tp, int up, int bytes)
{
    if (e->ntr == 4) launch_two_n<T, STAGE, HV, 4, kTPW<T>>(e, a, tp, up, bytes);
    else launch_two_n<T, STAGE, HV, 1, 1>(e, a, tp, up, bytes);
}

template <typename T>
int launch_mtile2(serbf_engine *e, int stage, int hv, StageArgs<T> a, const MTilePlan &tp)
{
    const int up = (e->tile_umax + 15) & ~15;
    const int bytes = mtile2_lds_bytes<T>(up, hv != 0, e->ntr);
    SERBF_CHECK(bytes <= 160 * 1024, SERBF_ERR_UNSUPPORTED, "matrix-core stage kernel: tile union too large");
    switch (stage * 2 + (hv ? 1 : 0)) {
    case 0: launch_two<T, 0, 0>(e, a, tp, up, bytes); break;
    case 1: launch_two<T, 0, 1>(e, a, tp, up, bytes); break;
    case 2: launch_two<T, 1, 0>(e, a, tp, up, bytes); break;
    case 3: launch_two<T, 1, 1>(e, a, tp, up, bytes); break;
    case 4: launch_two<T, 2, 0>(e, a, tp, up, bytes); break;
    case 5: launch_two<T, 2, 1>(e, a, tp, up, bytes); break;
    case 6: launch_two<T, 3, 0>(e, a, tp, up, bytes); break;
    case 7: launch_two<T, 3, 1>(e, a, tp, up, bytes); break;
    case 8: launch_two<T, 4, 0>(e, a, tp, up, bytes); break;
    default: launch_two<T, 4, 1>(e, a, tp, up, bytes); break;
    }
    SERBF_HIP(hipGetLastError());
    return SERBF_OK;
}

}  // namespace

// The matrix-core stage kernels serve 4 tracers (version 5 also 1 tracer) of 64 levels, flows
// whose horizontal velocity does not depend on the level (config mode), and tile plans whose
// unions fit (FP32 and FP64).
bool mtile_supported(const serbf_engine *e)
{
    return (e->ntr == 4 || e->ntr == 1) && e->S == 64 && e->cfg.mode == SERBF_MODE_CONFIG &&
           (e->cfg.flow == SERBF_FLOW_DEFORM || e->cfg.flow == SERBF_FLOW_SOLID_BODY) && e->have_plan &&
           e->tile_umax <= kMTileUcap && (size_t)e->N * e->ntr * e->S < (size_t)INT_MAX;
}

// hv: 0 none, 1 hyperviscosity in this stage, 2 split (stage 1 only; not handled here)
template <typename T>
int launch_mtile(serbf_engine *e, int stage, int hv, StageArgs<T> a)
{
    SERBF_CHECK(hv != 2, SERBF_ERR_UNSUPPORTED, "matrix-core stage kernel: split hyperviscosity");
    MTilePlan tp;
    tp.uni = e->d_tuni;
    tp.ucount = e->d_tcount;
    tp.lidx = e->d_lidx;
    tp.ustride = e->tile_ustride;
    a.nblk = e->ntiles;
    if constexpr (std::is_same<T, float>::value)
        if (e->kver == 7 && mvec_supported(e)) return launch_mvec(e, stage, hv, a, tp);
    return launch_mtile2<T>(e, stage, hv, a, tp);
}
template int launch_mtile<float>(serbf_engine *, int, int, StageArgs<float>);
template int launch_mtile<double>(serbf_engine *, int, int, StageArgs<double>);

}  // namespace serbf
