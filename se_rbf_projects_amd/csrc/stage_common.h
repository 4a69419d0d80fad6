// Shared device-side pieces of the stage kernels (rk4.hip, tile.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "engine.h"

namespace serbf {

enum VelKind { VEL_TC2 = SERBF_FLOW_TC2, VEL_SOLID = SERBF_FLOW_SOLID_BODY,
               VEL_DEFORM = SERBF_FLOW_DEFORM, VEL_ARRAY = 3 };

template <typename T>
struct StageArgs {
    const T *qin;       // gathered stage input     [N][ntr][LANES]
    const T *qT;        // step-start state          (stages 2-4)
    T *qout;            // stage output state        (stages 1-3), q_T (stage 4)
    T *F;               // accumulator / RHS output
    const int *idx;     // [N][n]
    const T *Wx, *Wy, *WL;
    const T *basis;     // [N][kBasis]
    const T *lev;       // [4][LANES]
    const int *vstart;  // [LANES]
    const T *vW;        // [LANES][kMaxNv]
    const T *tscal;     // time table
    const int *base;    // step counter
    const T *uarr;      // VEL_ARRAY: Up [N][3][LANES]
    const T *gamma;     // per-point hyperviscosity coefficient (hv_mode 1)
    const _Float16 *s16;  // per-point multiplier of hv_coef in FP16 (hv_mode 2; FP32 tile kernel)
    int N, n, Nv, nv;
    int flow, ref_mode, hv_mode;
    int step_off;
    int nblk;           // grid size (for the XCD remap)
    T dt, hdt, dt6, hv_coef, alpha;
};

__device__ __forceinline__ int xcd_remap(int b, int nblk)
{
    // Hardware dispatches work-group b to XCD b % 8.  Give XCD x the contiguous
    // block range [x*q + min(x,r), ...) so each XCD's L2 serves one region.
    const int q = nblk >> 3, r = nblk & 7;
    const int x = b & 7, i = b >> 3;
    return x * q + (x < r ? x : r) + i;
}

template <typename T>
struct Vec4 {
    T v[4];
};

template <typename T>
__device__ __forceinline__ Vec4<T> load4(const T *p)
{
    Vec4<T> r;
    if constexpr (sizeof(T) == 4) {
        const float4 x = *(const float4 *)p;
        r.v[0] = x.x; r.v[1] = x.y; r.v[2] = x.z; r.v[3] = x.w;
    } else {
        const double2 x = *(const double2 *)p, y = *(const double2 *)(p + 2);
        r.v[0] = x.x; r.v[1] = x.y; r.v[2] = y.x; r.v[3] = y.y;
    }
    return r;
}

template <typename T>
__device__ __forceinline__ void store4(T *p, const Vec4<T> &r)
{
    if constexpr (sizeof(T) == 4) {
        *(float4 *)p = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
    } else {
        *(double2 *)p = make_double2(r.v[0], r.v[1]);
        *(double2 *)(p + 2) = make_double2(r.v[2], r.v[3]);
    }
}

// stencil entry in LDS: (weight a, weight b, Delta^k weight, neighbour index)
template <typename T>
struct alignas(4 * sizeof(T)) MetaE {
    T a, b, l, idx_bits;
};

template <typename T>
__device__ __forceinline__ T idx_to_bits(int i)
{
    if constexpr (sizeof(T) == 4) return __int_as_float(i);
    else return __longlong_as_double((long long)i);
}
template <typename T>
__device__ __forceinline__ int bits_to_idx(T b)
{
    if constexpr (sizeof(T) == 4) return __float_as_int(b);
    else return (int)__double_as_longlong(b);
}

// horizontal velocity of a node at a level-independent flow (config mode)
template <typename T>
__device__ __forceinline__ void node_velocity(const StageArgs<T> &a, const T *B, const T *ts, T &upx, T &upy)
{
    if (a.flow == VEL_DEFORM) {
        const T c = ts[0], C2 = ts[1], S2 = ts[2];
        const T cl = T(0.5) - T(0.5) * (B[9] * C2 + B[10] * S2);
        const T cp = B[10] * C2 - B[9] * S2;
        upx = B[0] + c * (cl * B[3] + cp * B[6]);
        upy = B[1] + c * (cl * B[4] + cp * B[7]);
    } else {
        upx = B[0];
        upy = B[1];
    }
}

// velocity at level v (any flow)
template <typename T, int S>
__device__ __forceinline__ void level_velocity(const StageArgs<T> &a, const T *B, const T *ts, int pc, int v, T &upx,
                                               T &upy, T &upz)
{
    if (a.flow == VEL_TC2) {
        const T wt = ts[0];
        const T ga = wt * a.lev[S + v], gb = wt * a.lev[2 * S + v];
        upx = B[0] + (ga * B[3] + gb * B[6]);
        upy = B[1] + (ga * B[4] + gb * B[7]);
        upz = a.ref_mode ? B[2] + (ga * B[5] + gb * B[8]) : gb * B[9];
    } else if (a.flow == VEL_DEFORM) {
        const T c = ts[0], C2 = ts[1], S2 = ts[2];
        const T cl = T(0.5) - T(0.5) * (B[9] * C2 + B[10] * S2);
        const T cp = B[10] * C2 - B[9] * S2;
        const T w = c * a.lev[S + v];
        upx = B[0] + c * (cl * B[3] + cp * B[6]) + w * B[11];
        upy = B[1] + c * (cl * B[4] + cp * B[7]) + w * B[12];
        upz = a.ref_mode ? B[2] + c * (cl * B[5] + cp * B[8]) + w * B[13] : w;
    } else if (a.flow == VEL_SOLID) {
        upx = B[0];
        upy = B[1];
        upz = a.ref_mode ? B[2] : T(0);
    } else {
        const T *u = a.uarr + (size_t)pc * 3 * S + v;
        upx = u[0];
        upy = u[S];
        upz = u[2 * S];
    }
}


}  // namespace serbf
