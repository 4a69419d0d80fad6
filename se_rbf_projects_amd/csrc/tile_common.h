// Device pieces shared by the 16-node tile kernels (mtile.hip stage kernels, closure.hip
// evaluation): the tile plan, DPP row shifts, scalar-base 16-byte accesses, the MFMA row of a
// tile node, and the per-level vertical coefficient table over a lane's 12-level window.
#pragma once
#include <hip/hip_runtime.h>

#include "engine.h"

namespace serbf {

constexpr int kTileNodes16 = 16;

struct MTilePlan {
    const int *uni;              // [ntiles][ustride] global ids of the union (the tile's own nodes first)
    const int *ucount;           // [ntiles]
    const unsigned short *lidx;  // [N][n] union position of each stencil entry
    int ustride;
};

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

// MFMA operand layout (16x16x4, lane l, c = l % 16, g = l / 16): A (16 x 4) A[c][g], B (4 x 16)
// B[g][c]; result value i of lane (c, g) is D[4g + i][c] (FP32) or D[g + 4i][c] (FP64)
__device__ __forceinline__ v4f mfma4(float a, float b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4d mfma4(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// the value of lane c - 1 (c + 1) of the same 16-lane row, 0 for c = 0 (c = 15)
__device__ __forceinline__ float row_prev(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xF, 0xF, false));
}
__device__ __forceinline__ float row_next(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x101, 0xF, 0xF, false));
}

// 16-byte accesses at a byte offset below 4 GiB from a kernel-argument base: the scalar base plus
// a 32-bit offset register (global saddr form) instead of a 64-bit address per lane
template <typename V, typename P>
__device__ __forceinline__ V ld_at(const P *base, unsigned byte_off)
{
    return *(const V *)((const char *)base + byte_off);
}
template <typename V, typename P>
__device__ __forceinline__ void st_at(P *base, unsigned byte_off, const V &v)
{
    *(V *)((char *)base + byte_off) = v;
}

// Streaming forms for data a stage touches once (the RK4 update's inputs and outputs): the
// non-temporal hint keeps them from displacing the union columns, which the neighbouring tiles
// read again, from L2.
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 ld_nt4(const float *base, unsigned byte_off)
{
    const v4f v = __builtin_nontemporal_load((const v4f *)((const char *)base + byte_off));
    return make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ ushort4 ld_nt_h4(const _Float16 *base, unsigned byte_off)
{
    return __builtin_bit_cast(ushort4, __builtin_nontemporal_load((const v2u *)((const char *)base + byte_off)));
}
__device__ __forceinline__ void st_nt4(float *base, unsigned byte_off, float4 x)
{
    __builtin_nontemporal_store(v4f{x.x, x.y, x.z, x.w}, (v4f *)((char *)base + byte_off));
}

// MFMA row of tile node t: node g + 4i lands in row 4g + i, i.e. in result value i of row group g
__device__ __forceinline__ int vec_row(int t) { return 4 * (t & 3) + (t >> 2); }

// Vertical stencil over a lane's window of levels 4c - 4 .. 4c + 7 (lane c holds levels 4c .. 4c + 3
// of a 64-level column): Ct[l][r] is the weight of window position r in the stencil of level l,
// zero outside the stencil and for levels l >= Nv.
constexpr int kVecWin = 12;
__device__ __forceinline__ void build_vertical_table(float *Ct, int Nv, int nv, const int *vstart, const float *vW,
                                                     int tid, int nth)
{
    for (int e = tid; e < kMaxLevels * kVecWin; e += nth) {
        const int l = e / kVecWin, r = e - l * kVecWin;
        float cv = 0.f;
        if (l < Nv) {
            const int m = r - (vstart[l] - 4 * (l >> 2) + 4);
            if (m >= 0 && m < nv) cv = vW[l * kMaxNv + m];
        }
        Ct[e] = cv;
    }
}

// the vertical stencil of levels v0 .. v0 + 3 (v0 = 4c) of one node from its own column there (ov,
// lane c) and the lanes c -/+ 1 beside it: the window levels 4c - 4 .. 4c + 7 against the table rows
__device__ __forceinline__ void vertical4(const float *Ct, int v0, const float4 &ov, float z[4])
{
    const float o[4] = {ov.x, ov.y, ov.z, ov.w};
    float win[kVecWin];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        win[k] = row_prev(o[k]);
        win[4 + k] = o[k];
        win[8 + k] = row_next(o[k]);
    }
#pragma unroll
    for (int lb = 0; lb < 4; lb++) {
        const float4 *cr = (const float4 *)(Ct + (v0 + lb) * kVecWin);
        const float4 c0 = cr[0], c1 = cr[1], c2 = cr[2];
        const float cf[kVecWin] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x, c2.y, c2.z, c2.w};
        float zv = 0.f;
#pragma unroll
        for (int r = 0; r < kVecWin; r++) zv = fmaf(cf[r], win[r], zv);
        z[lb] = zv;
    }
}

// the vertical window of the tile kernels holds every stencil tap: level l's taps lie within
// levels 4 (l / 4) - 4 .. 4 (l / 4) + 7 (n_v <= 7 centred, one-sided at the ends)
inline bool vertical_window_fits(const serbf_engine *e)
{
    if (e->cfg.Nv > kMaxLevels || e->nv > kMaxNv) return false;
    for (int l = 0; l < e->cfg.Nv; l++) {
        const int b = e->h_vstart[l] - 4 * (l >> 2) + 4;
        if (b < 0 || b + e->nv > kVecWin) return false;
    }
    return true;
}

}  // namespace serbf
