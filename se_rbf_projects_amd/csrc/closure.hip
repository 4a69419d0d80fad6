// Learned hyperviscosity closure (gfx950): per-point features, the 10 -> 64 -> 64 -> 1 MLP
// (ReLU hidden layers) in BF16 on the matrix cores (v_mfma_f32_16x16x16_bf16, FP32
// accumulation), its backward pass, Adam, and the per-node stability bound of the multiplier.
//
// MFMA operand layout (16x16x16, lane l, c = l % 16, g = l / 16):
//   A (M x K): A[c][4g + j]      B (K x N): B[4g + j][c]      D (M x N): D[4g + i][c]
// Layer 1 is computed as Z1 = W1 X (M = hidden unit, N = sample), so its output tile is
// already the A operand of layer 2, Z2^T = H1^T W2^T (M = sample, N = unit).  The backward
// pass needs three transposes (H1, dZ2, dZ1), staged through the wavefront's LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "engine.h"
#include "stage_common.h"
#include "tile_common.h"

struct serbf_closure {
    int device = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    unsigned short *w1b = nullptr, *w2b = nullptr, *w2tb = nullptr;
    float *norm = nullptr;
    int step = 0;
    bool have_params = false;
    long n = 0, cap = 0;
    float *feat = nullptr, *A = nullptr, *R = nullptr;
    float *scap = nullptr;        // per-sample clamp of the training loss (optional)
    float *partials = nullptr, *grad = nullptr, *losses = nullptr;
    int loss_cap = 0;
    float *out = nullptr;
    long out_cap = 0;
    double s_max = 16.0;   // clamp of the multiplier in the training loss
};

namespace serbf {
namespace {

constexpr int kF = SERBF_CLOSURE_FEATURES;
constexpr int kH = SERBF_CLOSURE_HIDDEN;
constexpr int kP = SERBF_CLOSURE_PARAMS;
constexpr int kPS = 4936;   // partial-gradient record: parameters, loss, padding
constexpr int oB1 = kF * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kH, oLoss = kP;
constexpr int kMaxWaves = 1024;
constexpr float kKappa = 8.0f;
static_assert(oB3 + 1 == kP, "parameter layout");

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// round to nearest even
__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
typedef float fl2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
struct Bf2x2 {
    bf2 lo, hi;
};
// four floats to BF16, round to nearest even (v_cvt_pk_bf16_f32, two per instruction)
__device__ __forceinline__ v4s pack4(float a, float b, float c, float d)
{
    const Bf2x2 t = {__builtin_convertvector((fl2){a, b}, bf2), __builtin_convertvector((fl2){c, d}, bf2)};
    return __builtin_bit_cast(v4s, t);
}
__device__ __forceinline__ v4s ld4(const unsigned short *p) { return *(const v4s *)p; }
__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ float softplus_f(float o) { return o > 20.f ? o : log1pf(expf(o)); }
// true when a point's features hold a NaN or an infinity: the network would not pass them on
// (ReLU maps NaN to 0 or keeps it by its sign bit) but return an arbitrary finite output, so
// such a point takes the fallback s = 1
__device__ __forceinline__ bool bad_features(const float (&f)[SERBF_CLOSURE_FEATURES])
{
    float sum = 0.f;
#pragma unroll
    for (int k = 0; k < SERBF_CLOSURE_FEATURES; k++) sum += fabsf(f[k]);
    return !(sum <= 3.0e38f);
}
__device__ __forceinline__ float sigmoid_f(float o) { return 1.f / (1.f + expf(-o)); }
// hidden-layer activation: ReLU as a signed-integer maximum of the bits, one full-rate
// instruction (fmaxf(x, 0) takes two: IEEE mode makes it quiet a signalling NaN first).  Negative
// values and -0 are negative integers and give +0; positive values and positive NaNs pass.
__device__ __forceinline__ float relu(float x) { return __int_as_float(max(__float_as_int(x), 0)); }

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

struct NetView {
    const float *params;
    const unsigned short *w1b;    // W1 [64][16] (features 10..15 zero)
    const unsigned short *w2b;    // W2 [64][64], row = output unit
    const unsigned short *w2tb;   // W2^T [64][64]
    const float *norm;            // shift[16], scale[16]
};

NetView view(const serbf_closure *c) { return NetView{c->params, c->w1b, c->w2b, c->w2tb, c->norm}; }

// forward-pass operands held in registers by every lane
struct FwdRegs {
    v4s w1[4];
    v4s w2[4][4];
    float b1[4][4];
    float b2[4], w3[4];
    float b3;
};

__device__ __forceinline__ void load_fwd(const NetView &nv, int lane, FwdRegs &r)
{
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; t++) r.w1[t] = ld4(nv.w1b + (16 * t + c) * 16 + 4 * g);
#pragma unroll
    for (int kk = 0; kk < 4; kk++)
#pragma unroll
        for (int t = 0; t < 4; t++) r.w2[kk][t] = ld4(nv.w2b + (16 * t + c) * kH + 16 * kk + 4 * g);
#pragma unroll
    for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int i = 0; i < 4; i++) r.b1[t][i] = nv.params[oB1 + 16 * t + 4 * g + i];
        r.b2[t] = nv.params[oB2 + 16 * t + c];
        r.w3[t] = nv.params[oW3 + 16 * t + c];
    }
    r.b3 = nv.params[oB3];
}

// One tile of 16 samples.  xb = X[feature 4g + j][sample c].  Out: h1[t][i] = H1[unit 16t + 4g + i]
// [sample c], h2[t][i] = H2[sample 4g + i][unit 16t + c], o[i] = output of sample 4g + i.
__device__ __forceinline__ void fwd_tile(const FwdRegs &r, v4s xb, float h1[4][4], float h2[4][4], float o[4])
{
    // the biases are the accumulators' initial values
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const v4f z = mfma(r.w1[t], xb, v4f{r.b1[t][0], r.b1[t][1], r.b1[t][2], r.b1[t][3]});
#pragma unroll
        for (int i = 0; i < 4; i++) h1[t][i] = relu(z[i]);
    }
    v4f z2[4];
#pragma unroll
    for (int t = 0; t < 4; t++) z2[t] = v4f{r.b2[t], r.b2[t], r.b2[t], r.b2[t]};
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
        const v4s a = pack4(h1[kk][0], h1[kk][1], h1[kk][2], h1[kk][3]);
#pragma unroll
        for (int t = 0; t < 4; t++) z2[t] = mfma(a, r.w2[kk][t], z2[t]);
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) h2[t][i] = relu(z2[t][i]);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float s = 0.f;
#pragma unroll
        for (int t = 0; t < 4; t++) s = fmaf(r.w3[t], h2[t][i], s);
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) s += __shfl_xor(s, m, 64);
        o[i] = s + r.b3;
    }
}

// ---- the forward pass of the evaluation kernels: every layer as W H (M = unit, N = sample), so
// that a layer's result tile is already the next layer's B operand and the output reduction is two
// cross-row shuffles; layer 2 on v_mfma_f32_16x16x32_bf16.  Sample c of a tile lives in lanes
// (c, g), g = 0..3, which all end with its output.
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v4f mfma32(v8s a, v8s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0,
                                                    0);
}

// Layer 2's K block kk (32 units) in lane row g holds units 32 kk + 4g + i (i = 0..3) and
// 32 kk + 16 + 4g + i: the layer-1 result tiles 2kk and 2kk + 1 of that lane, packed as they are.
struct InferRegs {
    v4s w1[4];        // W1[16t + c][4g .. 4g + 3]
    v8s w2[2][4];     // W2[16tm + c][units of K block kk in row g]
    float b1[4][4];   // b1[16t + 4g + i]
    float b2[4][4];   // b2[16tm + 4g + i]
    float w3[4][4];   // w3[16tm + 4g + i]
    float b3;
};

__device__ __forceinline__ void load_infer(const NetView &nv, int lane, InferRegs &r)
{
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; t++) r.w1[t] = ld4(nv.w1b + (16 * t + c) * 16 + 4 * g);
#pragma unroll
    for (int kk = 0; kk < 2; kk++)
#pragma unroll
        for (int tm = 0; tm < 4; tm++) {
            const unsigned short *row = nv.w2b + (16 * tm + c) * kH + 32 * kk + 4 * g;
            const v4s lo = ld4(row), hi = ld4(row + 16);
            r.w2[kk][tm] = v8s{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            r.b1[t][i] = nv.params[oB1 + 16 * t + 4 * g + i];
            r.b2[t][i] = nv.params[oB2 + 16 * t + 4 * g + i];
            r.w3[t][i] = nv.params[oW3 + 16 * t + 4 * g + i];
        }
    r.b3 = nv.params[oB3];
}

// xb = X[feature 4g + j][sample c]; returns the output of sample c
__device__ __forceinline__ float fwd_infer(const InferRegs &r, v4s xb)
{
    // layer 1, ReLU applied to the BF16 pairs as a 16-bit signed maximum (rounding to BF16 and
    // ReLU commute: a negative value rounds to a negative BF16 or -0)
    v4s h1[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const v4f z = mfma(r.w1[t], xb, v4f{r.b1[t][0], r.b1[t][1], r.b1[t][2], r.b1[t][3]});
        const v4s zb = pack4(z[0], z[1], z[2], z[3]);
        h1[t] = __builtin_elementwise_max(zb, v4s{0, 0, 0, 0});
    }
    v8s hp[2];
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
        const v4s lo = h1[2 * kk], hi = h1[2 * kk + 1];
        hp[kk] = v8s{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    }
    float part = 0.f;
#pragma unroll
    for (int tm = 0; tm < 4; tm++) {
        v4f z2 = v4f{r.b2[tm][0], r.b2[tm][1], r.b2[tm][2], r.b2[tm][3]};
#pragma unroll
        for (int kk = 0; kk < 2; kk++) z2 = mfma32(r.w2[kk][tm], hp[kk], z2);
#pragma unroll
        for (int i = 0; i < 4; i++) part = fmaf(r.w3[tm][i], relu(z2[i]), part);
    }
    // sum over the four lane rows: row pairs by the 16-lane swap, then halves by the 32-lane swap
    // (each lane adds the same two operands as a xor-16 / xor-32 shuffle would)
    const auto p16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(part), __float_as_uint(part), false, false);
    part = __uint_as_float(p16[0]) + __uint_as_float(p16[1]);
    const auto p32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(part), __float_as_uint(part), false, false);
    part = __uint_as_float(p32[0]) + __uint_as_float(p32[1]);
    return part + r.b3;
}

// asinh(x) = log(|x| + sqrt(x^2 + 1)) with the sign of x, on the hardware log2 and square root;
// above 2^12, x^2 + 1 rounds to x^2 and the form log(2|x|) is used.  Its error is a few units of
// the last place of the result, except near 0, where forming 1 + |x| costs relative precision
// but the absolute error stays below about 1e-7: either is far below the BF16 rounding of the
// normalised feature it feeds.
__device__ __forceinline__ float asinh_fast(float x)
{
    const float ax = fabsf(x);
    const float y = ax > 4096.f ? 2.f * ax : ax + __builtin_sqrtf(fmaf(ax, ax, 1.f));
    return copysignf(__builtin_amdgcn_logf(y) * 0.69314718055994531f, x);
}

__device__ __forceinline__ float pick4(const float o[4], int i)
{
    return i == 0 ? o[0] : (i == 1 ? o[1] : (i == 2 ? o[2] : o[3]));
}

// ---------------------------------------------------------------------------
// features

template <typename T>
struct FeatArgs {
    StageArgs<T> s;          // state (qin), stencil, weights, basis, levels, vertical table, times
    const int *origin;       // step index of the run's first step within the closure cadence
    const T *WL1, *WL2;
    const float *hloc;
    const float *sbound;     // per-node stability bound of the multiplier, or null
    int ntr;
    float hbar, umax, R, dh;
};

template <typename T, int S>
__device__ void point_features(const FeatArgs<T> &a, int p, int tr, int v, float f[kF], T &l4out)
{
    const StageArgs<T> &s = a.s;
    const int NT = a.ntr;
    const int vv = v < s.Nv ? v : s.Nv - 1;
    T gx = T(0), gy = T(0), l1 = T(0), l2 = T(0), l4 = T(0);
    const size_t row = (size_t)p * s.n;
    const T *qcol = s.qin + (size_t)tr * S + vv;
    // batches of 8 neighbours: the index and value loads of a batch are all in flight before
    // the first use (the loop is memory-latency bound otherwise)
    constexpr int UB = 8;
    int k = 0;
    for (; k + UB <= s.n; k += UB) {
        int nb[UB];
        T q[UB];
#pragma unroll
        for (int u = 0; u < UB; u++) nb[u] = s.idx[row + k + u];
#pragma unroll
        for (int u = 0; u < UB; u++) q[u] = qcol[(size_t)nb[u] * NT * S];
#pragma unroll
        for (int u = 0; u < UB; u++) {
            const size_t o = row + k + u;
            gx = fma(s.Wx[o], q[u], gx);
            gy = fma(s.Wy[o], q[u], gy);
            l1 = fma(a.WL1[o], q[u], l1);
            l2 = fma(a.WL2[o], q[u], l2);
            l4 = fma(s.WL[o], q[u], l4);
        }
    }
    for (; k < s.n; k++) {
        const T q = qcol[(size_t)s.idx[row + k] * NT * S];
        gx = fma(s.Wx[row + k], q, gx);
        gy = fma(s.Wy[row + k], q, gy);
        l1 = fma(a.WL1[row + k], q, l1);
        l2 = fma(a.WL2[row + k], q, l2);
        l4 = fma(s.WL[row + k], q, l4);
    }
    const T *own = s.qin + ((size_t)p * NT + tr) * S;
    T dz = T(0);
    const int st = s.vstart[vv];
    for (int j = 0; j < s.nv; j++) {
        const int k = st + j;
        if (k >= 0 && k < s.Nv) dz = fma(s.vW[vv * kMaxNv + j], own[k], dz);
    }
    const T *ts = s.tscal + ((size_t)(*s.base + s.step_off) * 4) * kTimeScalars;
    T ux, uy, uz;
    level_velocity<T, S>(s, s.basis + (size_t)p * kBasis, ts, p, vv, ux, uy, uz);
    const T speed = sqrt(ux * ux + uy * uy + uz * uz);
    const T h = (T)a.hloc[p];
    const T h2 = h * h, h4 = h2 * h2, h8 = h4 * h4;
    const T K = (T)kKappa;
    f[0] = (float)asinh(K * h * sqrt(gx * gx + gy * gy));
    f[1] = (float)asinh(K * h2 * l1);
    f[2] = (float)asinh(K * h4 * l2);
    f[3] = (float)asinh(K * h8 * l4);
    f[4] = (float)(h / (T)a.hbar);
    f[5] = (float)(speed * s.dt / ((T)a.R * h));
    f[6] = (float)(speed / (T)a.umax);
    f[7] = (float)((T(vv) + T(0.5)) / T(s.Nv));
    f[8] = (float)own[vv];
    f[9] = (float)asinh(K * (T)a.dh * fabs(dz));
    l4out = l4;
}

// feat[(col * Nv + v) * 16 + k], lap[col * Nv + v]; a wavefront holds 64 / S columns
template <typename T, int S>
__global__ void __launch_bounds__(256) features_kernel(FeatArgs<T> a, float *feat, T *lap)
{
    constexpr int CPW = 64 / S;
    const int lane = threadIdx.x & 63;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long col = wave * CPW + lane / S;
    const int v = lane % S;
    const long ncol = (long)a.s.N * a.ntr;
    if (col >= ncol || v >= a.s.Nv) return;
    const int p = (int)(col / a.ntr), tr = (int)(col % a.ntr);
    float f[kF];
    T l4;
    point_features<T, S>(a, p, tr, v, f, l4);
    const size_t smp = (size_t)col * a.s.Nv + v;
#pragma unroll
    for (int k = 0; k < 16; k++) feat[smp * 16 + k] = k < kF ? f[k] : 0.f;
    lap[smp] = l4;
}

// closure coefficients: gamma[(p*ntr + tr)*S + v] = hv_coef * s, s = softplus(MLP(x)), with
// s clamped to s_max (and the node's bound), s = 1 where the features or s are not finite;
// levels >= Nv get 0.
//
// A wavefront pass covers NPW = 64/S nodes (lane = node slot x level) and a chunk of up to
// four tracers, i.e. 256 points.  The pass stages the nodes' stencil rows (indices and the
// five weight sets) in the wavefront's LDS with one coalesced load, then gathers 8 neighbours
// x 4 tracers per batch (32 independent loads in flight per lane), writes the normalised BF16
// features of its 256 points to LDS and runs them through the network as 16 tiles of 16.
template <typename T, int S>
struct ApplyGeom {
    static constexpr int NPW = 64 / S;
    static constexpr int TRC = 4;                       // tracers per pass
    static constexpr int PTS = 64 * TRC;                // points per pass
    static constexpr size_t XS = (size_t)PTS * 16 * 2;  // BF16 features
    static constexpr size_t OS = (size_t)PTS * 4;       // network outputs
    __host__ __device__ static size_t wave_bytes(int npad) { return XS + OS + (size_t)NPW * npad * (5 * sizeof(T) + 4); }
};

// 3 wavefronts per SIMD (168 registers and a small spill, against 2 wavefronts unconstrained):
// 12.2 ms per FP32 evaluation at 655362 x 60 x 4 instead of 21 ms
template <typename T, int S>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
apply_kernel(FeatArgs<T> a, NetView nv, T *gamma, _Float16 *s16, T hv_coef, float s_max, int every, int npad)
{
    using G = ApplyGeom<T, S>;
    constexpr int NPW = G::NPW, TRC = G::TRC, UB = 8;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    // re-evaluate only on steps that are multiples of `every` (the step counter lives on the
    // device, so one captured graph serves every cadence); other steps exit at once
    if ((*a.s.base + a.s.step_off + *a.origin) % every != 0) return;
    const StageArgs<T> &s = a.s;
    unsigned char *wb = lds_raw + (size_t)w * G::wave_bytes(npad);
    unsigned short *xs = (unsigned short *)wb;
    float *os = (float *)(wb + G::XS);
    T *mw = (T *)(wb + G::XS + G::OS);            // [5][NPW * npad]: Wx, Wy, Delta, Delta^2, Delta^k
    int *mi = (int *)(mw + 5 * NPW * npad);       // [NPW * npad]
    const int rows = NPW * npad;
    const int n = s.n, NT = a.ntr;
    const int slot = lane / S, v = lane % S;
    const int vv = v < s.Nv ? v : s.Nv - 1;
    const int tchunks = (NT + TRC - 1) / TRC;
    const long ngroups = ((long)s.N + NPW - 1) / NPW * tchunks;
    const T *ts = s.tscal + ((size_t)(*s.base + s.step_off) * 4) * kTimeScalars;
    const size_t cstride = (size_t)NT * S;
    const T K = (T)kKappa;
    // grid-stride over the passes; the loop bound is wavefront-uniform
    for (long grp = (long)blockIdx.x * 4 + w; grp < ngroups; grp += (long)gridDim.x * 4) {
        const long ng = grp / tchunks;
        const int tr0 = (int)(grp - ng * tchunks) * TRC;
        const int node0 = (int)(ng * NPW);
        for (int j = lane; j < rows; j += 64) {
            const int sl = j / npad, k = j - sl * npad;
            int p = node0 + sl;
            if (p >= s.N) p = s.N - 1;
            if (k < n) {
                const size_t o = (size_t)p * n + k;
                mi[j] = s.idx[o];
                mw[j] = s.Wx[o];
                mw[rows + j] = s.Wy[o];
                mw[2 * rows + j] = a.WL1[o];
                mw[3 * rows + j] = a.WL2[o];
                mw[4 * rows + j] = s.WL[o];
            } else {
                mi[j] = p;
#pragma unroll
                for (int m = 0; m < 5; m++) mw[m * rows + j] = T(0);
            }
        }
        wave_sync();
        int p = node0 + slot;
        const bool pok = p < s.N;
        if (!pok) p = s.N - 1;
        int trc[TRC];
#pragma unroll
        for (int t = 0; t < TRC; t++) trc[t] = tr0 + t < NT ? tr0 + t : NT - 1;
        T gx[TRC], gy[TRC], l1[TRC], l2[TRC], l4[TRC];
#pragma unroll
        for (int t = 0; t < TRC; t++) gx[t] = gy[t] = l1[t] = l2[t] = l4[t] = T(0);
        const int *ri = mi + slot * npad;
        const T *rw = mw + slot * npad;
        const T *qv = s.qin + vv;
        for (int k0 = 0; k0 < n; k0 += UB) {
            int nb[UB];
            T q[UB][TRC];
#pragma unroll
            for (int u = 0; u < UB; u++) nb[u] = ri[k0 + u];
#pragma unroll
            for (int u = 0; u < UB; u++)
#pragma unroll
                for (int t = 0; t < TRC; t++) q[u][t] = qv[(size_t)nb[u] * cstride + trc[t] * S];
#pragma unroll
            for (int u = 0; u < UB; u++) {
                const T wx = rw[k0 + u], wy = rw[rows + k0 + u], w1 = rw[2 * rows + k0 + u],
                        w2 = rw[3 * rows + k0 + u], wk = rw[4 * rows + k0 + u];
#pragma unroll
                for (int t = 0; t < TRC; t++) {
                    gx[t] = fma(wx, q[u][t], gx[t]);
                    gy[t] = fma(wy, q[u][t], gy[t]);
                    l1[t] = fma(w1, q[u][t], l1[t]);
                    l2[t] = fma(w2, q[u][t], l2[t]);
                    l4[t] = fma(wk, q[u][t], l4[t]);
                }
            }
        }
        // tracer-independent features
        T ux, uy, uz;
        level_velocity<T, S>(s, s.basis + (size_t)p * kBasis, ts, p, vv, ux, uy, uz);
        const T speed = sqrt(ux * ux + uy * uy + uz * uz);
        const T h = (T)a.hloc[p];
        const T h2 = h * h, h4 = h2 * h2, h8 = h4 * h4;
        float fc[4];
        fc[0] = (float)(h / (T)a.hbar);
        fc[1] = (float)(speed * s.dt / ((T)a.R * h));
        fc[2] = (float)(speed / (T)a.umax);
        fc[3] = (float)((T(vv) + T(0.5)) / T(s.Nv));
        const int vs0 = s.vstart[vv];
        bool bad[TRC];
#pragma unroll
        for (int t = 0; t < TRC; t++) {
            const T *own = s.qin + ((size_t)p * NT + trc[t]) * S;
            T dz = T(0);
            for (int j = 0; j < s.nv; j++) {
                const int k = vs0 + j;
                if (k >= 0 && k < s.Nv) dz = fma(s.vW[vv * kMaxNv + j], own[k], dz);
            }
            float f[kF];
            f[0] = (float)asinh(K * h * sqrt(gx[t] * gx[t] + gy[t] * gy[t]));
            f[1] = (float)asinh(K * h2 * l1[t]);
            f[2] = (float)asinh(K * h4 * l2[t]);
            f[3] = (float)asinh(K * h8 * l4[t]);
            f[4] = fc[0];
            f[5] = fc[1];
            f[6] = fc[2];
            f[7] = fc[3];
            f[8] = (float)own[vv];
            f[9] = (float)asinh(K * (T)a.dh * fabs(dz));
            bad[t] = bad_features(f);
            float xn[16];
#pragma unroll
            for (int k = 0; k < 16; k++) xn[k] = k < kF ? (f[k] - nv.norm[k]) * nv.norm[16 + k] : 0.f;
            v4s *xrow = (v4s *)(xs + (t * 64 + lane) * 16);
#pragma unroll
            for (int k = 0; k < 4; k++) xrow[k] = pack4(xn[4 * k], xn[4 * k + 1], xn[4 * k + 2], xn[4 * k + 3]);
        }
        wave_sync();
        // the network operands are (re)loaded after the features, so the two phases do not
        // hold their registers at the same time (cache hits after the first pass)
        InferRegs r;
        load_infer(nv, lane, r);
        const int tiles = 4 * (NT - tr0 < TRC ? NT - tr0 : TRC);
        // tile st holds the samples 16 st + c, i.e. tracer st / 4 of lanes (c, st % 4); every lane
        // keeps the outputs of its own samples
        float om[TRC];
#pragma unroll
        for (int t = 0; t < TRC; t++) om[t] = 0.f;
#pragma unroll 1
        for (int st = 0; st < tiles; st++) {
            const float o = fwd_infer(r, ld4(xs + (16 * st + c) * 16 + 4 * g));
            const bool mine = g == (st & 3);
#pragma unroll
            for (int t = 0; t < TRC; t++)
                if (mine && (st >> 2) == t) om[t] = o;
        }
#pragma unroll
        for (int t = 0; t < TRC; t++) {
            if (tr0 + t >= NT || !pok) continue;
            float sv = softplus_f(om[t]);
            if (bad[t] || !(sv >= 0.f)) sv = 1.f;   // non-finite features or output
            if (sv > s_max) sv = s_max;
            if (a.sbound && sv > a.sbound[p]) sv = a.sbound[p];
            if (v >= s.Nv) sv = 0.f;
            const size_t o = ((size_t)p * NT + tr0 + t) * S + v;
            if (s16) s16[o] = (_Float16)sv;
            else gamma[o] = hv_coef * (T)sv;
        }
        wave_sync();   // the LDS of the pass is rewritten by the next one
    }
}

// ---------------------------------------------------------------------------
// Evaluation on 16-node tiles (FP32, 4 tracers of 64 levels), the form the FP32 time loop uses.
// The five stencil operators of the features (Dx, Dy, Delta, Delta^2, Delta^k) are FP32 products
// of dense 16 x U weight blocks with the tile's union columns on v_mfma_f32_16x16x4f32, laid out as
// in the vector stage kernel: a wavefront per tracer, lane (c, g) loads levels 4c .. 4c + 3 of a
// union column, so one 16-byte load feeds 4 level blocks x 5 operators and the lane ends with nodes
// g + 4i at levels 4c .. 4c + 3.  The vertical derivative comes from the tile's own columns (the
// first 16 union positions) by DPP row shifts over a 12-level window.  The column gathers of the
// per-point form (55 neighbours x 5 operators of vector FMAs per point) made it vector-ALU bound.
// Then, per chunk of 64 points (one per lane), the features go through the wavefront's LDS into
// the network's operand layout and through the network as 4 tiles of 16.
constexpr int kCOps = 5;
constexpr int kNodeRec = 17;   // per tile node in LDS: velocity basis (16) and spacing h

typedef _Float16 v8h __attribute__((ext_vector_type(8)));

__device__ __forceinline__ v4f mfma32h(v8h a, v8h b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// x as hi + lo, two FP16 values (hi holds 11 significant bits, lo the next 11), saturated to the
// FP16 range (NaN stays NaN)
__device__ __forceinline__ void split16(float x, _Float16 &hi, _Float16 &lo)
{
    const float xs = x > 65504.f ? 65504.f : (x < -65504.f ? -65504.f : x);
    hi = (_Float16)xs;
    lo = (_Float16)(xs - (float)hi);
}

typedef __fp16 hp2 __attribute__((ext_vector_type(2)));

// x0, x1 (|x| <= 1, or non-finite) as packed FP16 pairs hi and lo with x = hi + lo + O(2^-22 |x|):
// hi rounded toward zero, the remainder x - hi exact in FP32, lo its FP16 rounding toward zero
__device__ __forceinline__ void split2(float x0, float x1, unsigned &hi, unsigned &lo)
{
    const hp2 h = __builtin_amdgcn_cvt_pkrtz(x0, x1);
    const hp2 l = __builtin_amdgcn_cvt_pkrtz(x0 - (float)h[0], x1 - (float)h[1]);
    hi = __builtin_bit_cast(unsigned, h);
    lo = __builtin_bit_cast(unsigned, l);
}

// softplus on the hardware exp2 / log2: max(o, 0) + log(1 + e^-|o|) (for e^-|o| below 2^-24 the
// second term rounds to 0, an absolute error below 6e-8)
__device__ __forceinline__ float softplus_fast(float o)
{
    const float e = __builtin_amdgcn_exp2f(-fabsf(o) * 1.4426950408889634f);
    return fmaxf(o, 0.f) + __builtin_amdgcn_logf(1.f + e) * 0.69314718055994531f;
}

// union positions padded to the 32 columns of a product step
__host__ __device__ inline int cvec_u32(int up) { return (up + 31) & ~31; }

__host__ __device__ inline size_t cvec_lds_bytes(int up)
{
    const size_t wld = (size_t)cvec_u32(up) + 8;               // FP16 row stride (16 bytes of padding)
    const size_t wbytes = 2 * kCOps * 16 * wld * 2;             // hi and lo weight blocks
    const size_t nbytes = 4 * (size_t)64 * 16 * 2;              // after the products: per wavefront the BF16
                                                                 // features of 64 points
    return (size_t)(16 * kNodeRec + kMaxLevels * kVecWin) * 4 + (size_t)cvec_u32(up) * 4 +
           (wbytes > nbytes ? wbytes : nbytes);
}

// Per node and tracer the maximum of |q| over the levels of the FP32 state [N][4][64] (the
// padding levels hold zeros), for the power-of-two scale of a tile's FP16 products: 16 lanes per
// (node, tracer), one 16-byte load each.  It runs on the evaluation steps of the cadence only, as
// the closure kernel.
__global__ void __launch_bounds__(256) nodemax_kernel(FeatArgs<float> a, float *nmax, int every)
{
    const StageArgs<float> &s = a.s;
    if ((*s.base + s.step_off + *a.origin) % every != 0) return;
    // grid stride over whole wavefronts (N * 64 is a multiple of 64): every lane of a wavefront
    // takes part in each shuffle
    const size_t n4 = (size_t)s.N * 64;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) {
        const float4 v = ((const float4 *)s.qin)[i];
        float m = fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        if ((threadIdx.x & 15) == 0) nmax[i >> 4] = m;
    }
}

// 2 wavefronts per SIMD (256 registers: 80 accumulators, the column loads of the next step and
// the split operands of this one, then the network)
template <bool S16>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
cvec_kernel(FeatArgs<float> a, NetView nv, MTilePlan tp, int up, const float *nmax, float *gamma, _Float16 *s16,
            float hv_coef, float s_max, int every)
{
    const StageArgs<float> &s = a.s;
    // re-evaluate only on steps that are multiples of `every` (as the per-point form)
    if ((*s.base + s.step_off + *a.origin) % every != 0) return;
    extern __shared__ __align__(16) unsigned char cv_lds[];
    constexpr int S = 64, NTR = 4, NTH = 256;
    const int wld = cvec_u32(up) + 8;
    float *nrec = (float *)cv_lds;                          // [16][kNodeRec]
    float *Ct = nrec + 16 * kNodeRec;                        // [64][kVecWin]
    int *uoff = (int *)(Ct + kMaxLevels * kVecWin);          // [u32] byte offsets of the union columns
    _Float16 *Wh = (_Float16 *)(uoff + cvec_u32(up));        // [kCOps][16][wld] weights, hi part
    _Float16 *Wl = Wh + kCOps * 16 * wld;                    // [kCOps][16][wld] lo part
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int tile = xcd_remap(blockIdx.x, s.nblk);
    const int p0 = tile * 16;
    const int U = tp.ucount[tile];
    const int u32 = (U + 31) & ~31;
    const int n = s.n;
    const float *ts = s.tscal + ((size_t)(*s.base + s.step_off) * 4) * kTimeScalars;
    const int *uni = tp.uni + (size_t)tile * tp.ustride;
    const unsigned lo4 = (unsigned)(w * S + 4 * c) * 4u;

    // ---- set-up: node records, union offsets, vertical table, dense weight blocks ----
    {
        const int t = tid >> 4, k = tid & 15;   // 256 threads: one basis value each
        const int p = p0 + t < s.N ? p0 + t : s.N - 1;
        nrec[t * kNodeRec + k] = s.basis[(size_t)p * kBasis + k];
        if (k == 0) nrec[t * kNodeRec + 16] = a.hloc[p];
    }
    for (int k = tid; k < u32; k += NTH) uoff[k] = (k < U ? uni[k] : uni[0]) * (NTR * S * 4);
    build_vertical_table(Ct, s.Nv, s.nv, s.vstart, s.vW, tid, NTH);
    for (int e = tid; e < kCOps * 16 * wld; e += NTH) ((unsigned *)Wh)[e] = 0u;   // hi and lo
    __syncthreads();
    {
        // the operators scaled to h, h, h^2, h^4, h^8 of their row's node (dimensionless, so
        // FP16 holds them: row sums of a few thousand at most on quasi-uniform node sets)
        const int ne = (p0 + 16 <= s.N ? 16 : s.N - p0) * n;
        const float rn = 1.f / (float)n;
        const float *src[kCOps] = {s.Wx, s.Wy, a.WL1, a.WL2, s.WL};
        for (int e = tid; e < ne; e += NTH) {
            // (e + 1/2) / n is at least 1/(2n) away from an integer, far above the rounding
            const int t = (int)(((float)e + 0.5f) * rn);
            const size_t o = (size_t)p0 * n + e;
            const int col = vec_row(t) * wld + tp.lidx[o];
            const float h = nrec[t * kNodeRec + 16], h2 = h * h, h4 = h2 * h2;
            const float hk[kCOps] = {h, h, h2, h4, h4 * h4};
#pragma unroll
            for (int op = 0; op < kCOps; op++) split16(src[op][o] * hk[op], Wh[op * 16 * wld + col], Wl[op * 16 * wld + col]);
        }
    }
    // The tracer's power-of-two scale over the tile's union, 2^-E with |q| <= 2^E (qm = f 2^E,
    // 1/2 <= f < 1): the products then keep about 22 bits relative to the largest value of each
    // sum, as FP32 sums do, and values below 2^-14 of it lose bits into FP16 subnormals.
    float qm = 0.f;
    for (int k = lane; k < U; k += 64) qm = fmaxf(qm, nmax[(size_t)uni[k] * 4 + w]);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) qm = fmaxf(qm, __shfl_xor(qm, o, 64));
    float sig = 1.f;
    if (qm > 0.f && qm <= 3.0e38f) {
        int E;
        (void)frexpf(qm, &E);
        // |E| <= 100 keeps sig and the K / sig that undoes it finite (maxima beyond 2^100 then
        // saturate the FP16 products, below 2^-100 they lose bits: far outside any tracer field)
        sig = ldexpf(1.f, -(E < -100 ? -100 : (E > 100 ? 100 : E)));
    }
    __syncthreads();

    // ---- D_op (16 nodes x 64 levels of tracer w) = W_op (16 x u32) x Q (u32 x 64) ----
    // Each product is hi(W) hi(Q) + hi(W) lo(Q) + lo(W) hi(Q) on v_mfma_f32_16x16x32_f16 with FP32
    // accumulation (about 22 significant bits per term).  Lane (c, g) holds the union columns
    // k0 + 8g .. k0 + 8g + 7 of a 32-column step: the A operand row c from LDS, the B operand
    // levels 4c + lb of the 8 columns from 8 16-byte loads.
    v4f acc[kCOps][4];
#pragma unroll
    for (int op = 0; op < kCOps; op++)
#pragma unroll
        for (int lb = 0; lb < 4; lb++) acc[op][lb] = v4f{0.f, 0.f, 0.f, 0.f};
    const _Float16 *wah = Wh + c * wld + 8 * g;
    const _Float16 *wal = Wl + c * wld + 8 * g;
    float4 b[8];
    auto load_step = [&](int k0) {
#pragma unroll
        for (int j = 0; j < 8; j++) b[j] = ld_at<float4>(s.qin, (unsigned)uoff[k0 + 8 * g + j] + lo4);
    };
    // the scaled values are at most 1 in magnitude: no saturation needed
    auto split_step = [&](v8h (&bh)[4], v8h (&bl)[4]) {
        unsigned uh[4][4], ul[4][4];   // [level block][column pair]
#pragma unroll
        for (int jp = 0; jp < 4; jp++) {
            const float4 x0 = b[2 * jp], x1 = b[2 * jp + 1];
            split2(x0.x * sig, x1.x * sig, uh[0][jp], ul[0][jp]);
            split2(x0.y * sig, x1.y * sig, uh[1][jp], ul[1][jp]);
            split2(x0.z * sig, x1.z * sig, uh[2][jp], ul[2][jp]);
            split2(x0.w * sig, x1.w * sig, uh[3][jp], ul[3][jp]);
        }
#pragma unroll
        for (int lb = 0; lb < 4; lb++) {
            bh[lb] = __builtin_bit_cast(v8h, uh[lb]);
            bl[lb] = __builtin_bit_cast(v8h, ul[lb]);
        }
    };
    auto step = [&](int k0, const v8h (&bh)[4], const v8h (&bl)[4]) {
#pragma unroll
        for (int op = 0; op < kCOps; op++) {
            const v8h ah = *(const v8h *)(wah + op * 16 * wld + k0);
            const v8h al = *(const v8h *)(wal + op * 16 * wld + k0);
#pragma unroll
            for (int lb = 0; lb < 4; lb++) {
                acc[op][lb] = mfma32h(al, bh[lb], acc[op][lb]);
                acc[op][lb] = mfma32h(ah, bl[lb], acc[op][lb]);
                acc[op][lb] = mfma32h(ah, bh[lb], acc[op][lb]);
            }
        }
    };
    load_step(0);
    for (int k0 = 0; k0 < u32; k0 += 32) {
        v8h bh[4], bl[4];
        split_step(bh, bl);
        if (k0 + 32 < u32) load_step(k0 + 32);
        __builtin_amdgcn_sched_barrier(0);
        step(k0, bh, bl);
    }
    // own columns after the products (they come from L2 by then): the vertical derivative and the
    // state feature (in a last tile past N the union's other columns, whose results are not stored)
    float4 own[4];
    if (p0 + 16 <= s.N) {
#pragma unroll
        for (int j = 0; j < 4; j++) own[j] = ld_at<float4>(s.qin, (unsigned)(p0 + 4 * j + g) * (NTR * S * 4u) + lo4);
    } else {
        const int nown = s.N - p0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = 4 * j + g;
            const int id = k < nown ? p0 + k : uni[k < U ? k : 0];
            own[j] = ld_at<float4>(s.qin, (unsigned)id * (NTR * S * 4u) + lo4);
        }
    }
    const int v0 = 4 * c;
    float dz[4][4];   // [node i][level block]
#pragma unroll
    for (int i = 0; i < 4; i++) vertical4(Ct, v0, own[i], dz[i]);
    __syncthreads();   // every wavefront is done with the weight blocks: the network staging reuses them

    // ---- features and the network, 64 points (one per lane) at a time ----
    unsigned short *xs = (unsigned short *)Wh + w * (64 * 16);            // [64][16] BF16
    InferRegs r;
    load_infer(nv, lane, r);
    const float K = kKappa, Kq = kKappa / sig;   // the products carry h^k and the scale sig
    const float ihbar = 1.f / a.hbar, iumax = 1.f / a.umax, inv = 1.f / (float)s.Nv;
#pragma unroll 1
    for (int i = 0; i < 4; i++) {
        const int t = g + 4 * i;
        const int pr = p0 + t;
        const int p = pr < s.N ? pr : s.N - 1;
        const float *B = nrec + t * kNodeRec;
        const float h = B[16];
        const float cfl = s.dt / (a.R * h);
        const float ov[4] = {own[i].x, own[i].y, own[i].z, own[i].w};
        float sv4[4];
#pragma unroll
        for (int lb = 0; lb < 4; lb++) {
            const int v = v0 + lb;
            const int vv = v < s.Nv ? v : s.Nv - 1;
            const float gx = acc[0][lb][i], gy = acc[1][lb][i];
            float ux, uy, uz;
            level_velocity<float, S>(s, B, ts, p, vv, ux, uy, uz);
            const float speed = sqrtf(ux * ux + uy * uy + uz * uz);
            float f[kF];
            f[0] = asinh_fast(Kq * sqrtf(gx * gx + gy * gy));
            f[1] = asinh_fast(Kq * acc[2][lb][i]);
            f[2] = asinh_fast(Kq * acc[3][lb][i]);
            f[3] = asinh_fast(Kq * acc[4][lb][i]);
            f[4] = h * ihbar;
            f[5] = speed * cfl;
            f[6] = speed * iumax;
            f[7] = ((float)vv + 0.5f) * inv;
            f[8] = ov[lb];
            f[9] = asinh_fast(K * a.dh * fabsf(dz[i][lb]));
            const bool bad = bad_features(f);
            float xn[16];
#pragma unroll
            for (int k = 0; k < 16; k++) xn[k] = k < kF ? (f[k] - nv.norm[k]) * nv.norm[16 + k] : 0.f;
            v4s *xrow = (v4s *)(xs + lane * 16);
#pragma unroll
            for (int k = 0; k < 4; k++) xrow[k] = pack4(xn[4 * k], xn[4 * k + 1], xn[4 * k + 2], xn[4 * k + 3]);
            wave_sync();
            // tile st holds the points of lanes (c, st); every lane keeps its own point's output
            float om = 0.f;
#pragma unroll 1
            for (int st = 0; st < 4; st++) {
                const float o = fwd_infer(r, ld4(xs + (16 * st + c) * 16 + 4 * g));
                if (g == st) om = o;
            }
            float sv = softplus_fast(om);
            if (bad || !(sv >= 0.f)) sv = 1.f;   // non-finite features or output
            if (sv > s_max) sv = s_max;
            if (a.sbound && sv > a.sbound[p]) sv = a.sbound[p];
            if (v >= s.Nv) sv = 0.f;
            sv4[lb] = sv;
            wave_sync();   // xs and os are rewritten by the next chunk
        }
        if (pr < s.N) {
            const unsigned m = (unsigned)p * (NTR * S) + (unsigned)(w * S + v0);
            if (S16) {
                ushort4 hv;
                hv.x = __builtin_bit_cast(unsigned short, (_Float16)sv4[0]);
                hv.y = __builtin_bit_cast(unsigned short, (_Float16)sv4[1]);
                hv.z = __builtin_bit_cast(unsigned short, (_Float16)sv4[2]);
                hv.w = __builtin_bit_cast(unsigned short, (_Float16)sv4[3]);
                st_at<ushort4>(s16, m * 2u, hv);
            } else {
                st_at<float4>(gamma, m * 4u, make_float4(hv_coef * sv4[0], hv_coef * sv4[1], hv_coef * sv4[2],
                                                          hv_coef * sv4[3]));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// training

// one wavefront per strided set of 16-sample tiles; writes its gradient record to partials
__global__ void __launch_bounds__(256) train_kernel(NetView nv, const float *feat, const float *A, const float *R,
                                                    const float *cap, long first, int ntiles, float inv_count,
                                                    float s_max, float *partials)
{
    __shared__ alignas(16) unsigned short w2t[kH * kH];
    __shared__ alignas(16) unsigned short h1s[4][kH * 16];
    __shared__ alignas(16) unsigned short dz2s[4][16 * kH];
    __shared__ alignas(16) unsigned short dz1s[4][kH * 16];
    __shared__ alignas(16) unsigned short xs[4][16 * 16];
    for (int k = threadIdx.x; k < kH * kH; k += blockDim.x) w2t[k] = nv.w2tb[k];
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int wave = blockIdx.x * 4 + w, nwaves = gridDim.x * 4;
    FwdRegs r;
    load_fwd(nv, lane, r);
    float nsh[4], nsc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        nsh[j] = nv.norm[4 * g + j];
        nsc[j] = (4 * g + j) < kF ? nv.norm[16 + 4 * g + j] : 0.f;
    }
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
    v4f dw1[4] = {zero, zero, zero, zero};
    v4f dw2[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) dw2[a][b] = zero;
    float db1[4][4] = {}, db2[4] = {}, dw3[4] = {};
    float db3 = 0.f, loss = 0.f;
    for (int tile = wave; tile < ntiles; tile += nwaves) {
        const long s0 = first + (long)tile * 16;
        const float4 fv = *(const float4 *)(feat + (s0 + c) * 16 + 4 * g);
        const float x0 = (fv.x - nsh[0]) * nsc[0], x1 = (fv.y - nsh[1]) * nsc[1];
        const float x2 = (fv.z - nsh[2]) * nsc[2], x3 = (fv.w - nsh[3]) * nsc[3];
        const v4s xb = pack4(x0, x1, x2, x3);
        *(v4s *)&xs[w][c * 16 + 4 * g] = xb;   // X[sample c][feature 4g + j]
        float h1[4][4], h2[4][4], o[4];
        fwd_tile(r, xb, h1, h2, o);
        float go[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const long si = s0 + 4 * g + i;
            const float av = A[si], rv = R[si];
            // s = min(softplus(o), s_max, cap): no gradient through the clamp
            const float smx = cap ? fminf(s_max, cap[si]) : s_max;
            const float sp = softplus_f(o[i]);
            const bool clamped = sp > smx;
            const float e = ((clamped ? smx : sp) - 1.f) * av - rv;
            if (c == 0) loss = fmaf(e, e, loss);
            go[i] = clamped ? 0.f : 2.f * e * av * sigmoid_f(o[i]) * inv_count;
        }
        if (c == 0) db3 += (go[0] + go[1]) + (go[2] + go[3]);
        float dz2[4][4];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            float acc = 0.f, bsum = 0.f;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                acc = fmaf(go[i], h2[t][i], acc);
                dz2[t][i] = h2[t][i] > 0.f ? go[i] * r.w3[t] : 0.f;
                bsum += dz2[t][i];
            }
            dw3[t] += acc;
            db2[t] += bsum;
        }
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                h1s[w][(16 * t + 4 * g + i) * 16 + c] = f2bf(h1[t][i]);      // H1[unit][sample]
                dz2s[w][(4 * g + i) * kH + 16 * t + c] = f2bf(dz2[t][i]);    // dZ2[sample][unit]
            }
        wave_sync();
        // dW2[n][h] += sum_s dZ2[s][n] H1[s][h]   (M = n, N = h, K = s)
#pragma unroll
        for (int tm = 0; tm < 4; tm++) {
            const v4s a = pack4(dz2[tm][0], dz2[tm][1], dz2[tm][2], dz2[tm][3]);
#pragma unroll
            for (int tn = 0; tn < 4; tn++) dw2[tm][tn] = mfma(a, ld4(&h1s[w][(16 * tn + c) * 16 + 4 * g]), dw2[tm][tn]);
        }
        // dH1^T[h][s] = sum_n W2^T[h][n] dZ2^T[n][s]   (M = h, N = s, K = n)
        float dz1[4][4];
#pragma unroll
        for (int tm = 0; tm < 4; tm++) {
            v4f acc = zero;
#pragma unroll
            for (int kk = 0; kk < 4; kk++)
                acc = mfma(ld4(&w2t[(16 * tm + c) * kH + 16 * kk + 4 * g]), ld4(&dz2s[w][c * kH + 16 * kk + 4 * g]),
                           acc);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                dz1[tm][i] = h1[tm][i] > 0.f ? acc[i] : 0.f;
                db1[tm][i] += dz1[tm][i];
                dz1s[w][(16 * tm + 4 * g + i) * 16 + c] = f2bf(dz1[tm][i]);   // dZ1[unit][sample]
            }
        }
        wave_sync();
        // dW1[h][f] += sum_s dZ1[s][h] X[s][f]   (M = h, N = f, K = s)
        v4s bx;
#pragma unroll
        for (int j = 0; j < 4; j++) bx[j] = (short)xs[w][(4 * g + j) * 16 + c];
#pragma unroll
        for (int tm = 0; tm < 4; tm++) dw1[tm] = mfma(ld4(&dz1s[w][(16 * tm + c) * 16 + 4 * g]), bx, dw1[tm]);
        wave_sync();
    }
    float *P = partials + (size_t)wave * kPS;
#pragma unroll
    for (int tm = 0; tm < 4; tm++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int h = 16 * tm + 4 * g + i;
            if (c < kF) P[h * kF + c] = dw1[tm][i];
#pragma unroll
            for (int tn = 0; tn < 4; tn++) P[oW2 + h * kH + 16 * tn + c] = dw2[tm][tn][i];
            float b = db1[tm][i];
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) b += __shfl_xor(b, m, 64);
            if (c == 0) P[oB1 + h] = b;
        }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        float a = db2[t], b = dw3[t];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        b += __shfl_xor(b, 16, 64);
        b += __shfl_xor(b, 32, 64);
        if (g == 0) {
            P[oB2 + 16 * t + c] = a;
            P[oW3 + 16 * t + c] = b;
        }
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        db3 += __shfl_xor(db3, m, 64);
        loss += __shfl_xor(loss, m, 64);
    }
    if (lane == 0) {
        P[oB3] = db3;
        P[oLoss] = loss;
    }
}

__device__ __forceinline__ void store_bf16_copies(int k, float p, unsigned short *w1b, unsigned short *w2b,
                                                  unsigned short *w2tb)
{
    if (k < oB1) {
        w1b[(k / kF) * 16 + k % kF] = f2bf(p);
    } else if (k >= oW2 && k < oB2) {
        const int q = k - oW2, n = q / kH, h = q % kH;
        w2b[n * kH + h] = f2bf(p);
        w2tb[h * kH + n] = f2bf(p);
    }
}

// sum of the wavefront records; optional Adam step on the parameters and their BF16 copies
__global__ void reduce_adam_kernel(float *params, float *m, float *v, unsigned short *w1b, unsigned short *w2b,
                                   unsigned short *w2tb, const float *partials, int nwaves, float *grad,
                                   float *losses, int loss_idx, float loss_scale, int update, float lr, float b1,
                                   float b2, float eps, float bc1, float bc2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > oLoss) return;
    float s = 0.f;
    for (int w = 0; w < nwaves; w++) s += partials[(size_t)w * kPS + k];
    if (k == oLoss) {
        s *= loss_scale;
        grad[k] = s;
        if (losses) losses[loss_idx] = s;
        return;
    }
    grad[k] = s;
    if (!update) return;
    const float mk = b1 * m[k] + (1.f - b1) * s;
    const float vk = b2 * v[k] + (1.f - b2) * s * s;
    m[k] = mk;
    v[k] = vk;
    const float p = params[k] - lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
    params[k] = p;
    store_bf16_copies(k, p, w1b, w2b, w2tb);
}

__global__ void pack_kernel(const float *params, unsigned short *w1b, unsigned short *w2b, unsigned short *w2tb)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kH * 16 && k % 16 >= kF) w1b[k] = 0;
    if (k < kP) store_bf16_copies(k, params[k], w1b, w2b, w2tb);
}

__global__ void __launch_bounds__(256) eval_kernel(NetView nv, const float *feat, long first, int ntiles, float *out)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    const int wave = blockIdx.x * 4 + w, nwaves = gridDim.x * 4;
    FwdRegs r;
    load_fwd(nv, lane, r);
    float nsh[4], nsc[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        nsh[j] = nv.norm[4 * g + j];
        nsc[j] = (4 * g + j) < kF ? nv.norm[16 + 4 * g + j] : 0.f;
    }
    for (int tile = wave; tile < ntiles; tile += nwaves) {
        const long s0 = first + (long)tile * 16;
        const float4 fv = *(const float4 *)(feat + (s0 + c) * 16 + 4 * g);
        const v4s xb = pack4((fv.x - nsh[0]) * nsc[0], (fv.y - nsh[1]) * nsc[1], (fv.z - nsh[2]) * nsc[2],
                             (fv.w - nsh[3]) * nsc[3]);
        float h1[4][4], h2[4][4], o[4];
        fwd_tile(r, xb, h1, h2, o);
        if (c < 4) out[s0 - first + 4 * g + c] = pick4(o, c);
    }
}

// nearest-neighbour chord of every node (stencil entry 1 is the nearest after the node itself)
__global__ void hloc_kernel(int N, int n, const double *x, const double *y, const double *z, const int *idx,
                            float *h)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    const int q = idx[(size_t)p * n + (n > 1 ? 1 : 0)];
    const double dx = x[q] - x[p], dy = y[q] - y[p], dz = z[q] - z[p];
    h[p] = (float)sqrt(dx * dx + dy * dy + dz * dz);
}

int launch_grid_waves(int ntiles)
{
    int waves = ntiles < kMaxWaves ? ntiles : kMaxWaves;
    if (waves < 4) waves = 4;
    return (waves + 3) / 4;   // blocks of 4 wavefronts
}

int run_grad(serbf_closure *c, long first, long count, bool update, float lr, int loss_idx)
{
    SERBF_CHECK(c->have_params, SERBF_ERR_STATE, "closure: parameters not set");
    SERBF_CHECK(count > 0 && count % 16 == 0 && first >= 0 && first + count <= c->n, SERBF_ERR_ARG,
                "closure: sample range must lie in the set and be a multiple of 16");
    const int ntiles = (int)(count / 16);
    const int blocks = launch_grid_waves(ntiles);
    hipLaunchKernelGGL(train_kernel, dim3(blocks), dim3(256), 0, c->stream, view(c), c->feat, c->A, c->R, c->scap,
                       first, ntiles, 1.f / (float)count, (float)c->s_max, c->partials);
    SERBF_HIP(hipGetLastError());
    float bc1 = 1.f, bc2 = 1.f;
    if (update) {
        c->step++;
        bc1 = 1.f - powf(0.9f, (float)c->step);
        bc2 = 1.f - powf(0.999f, (float)c->step);
    }
    hipLaunchKernelGGL(reduce_adam_kernel, dim3((kPS + 255) / 256), dim3(256), 0, c->stream, c->params, c->adam_m,
                       c->adam_v, c->w1b, c->w2b, c->w2tb, c->partials, blocks * 4, c->grad,
                       loss_idx >= 0 ? c->losses : nullptr, loss_idx, 1.f / (float)count, update ? 1 : 0, lr, 0.9f,
                       0.999f, 1e-8f, bc1, bc2);
    SERBF_HIP(hipGetLastError());
    return SERBF_OK;
}

template <typename T>
FeatArgs<T> feat_args(serbf_engine *e, StateBuffers<T> &sb, const T *q, double dt, const T *tscal, const int *base,
                      int step_off)
{
    FeatArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    StageArgs<T> &s = a.s;
    s.qin = q;
    s.idx = e->d_idx;
    s.Wx = sb.Wx; s.Wy = sb.Wy; s.WL = sb.WL;
    s.basis = sb.basis;
    s.lev = sb.lev;
    s.vstart = e->d_vstart;
    s.vW = sb.vW;
    s.tscal = tscal;
    s.base = base;
    s.step_off = step_off;
    a.origin = e->d_base + 1;   // zero, outside the time loop
    s.N = e->N; s.n = e->n; s.Nv = e->cfg.Nv; s.nv = e->nv;
    s.flow = e->cfg.flow;
    s.ref_mode = e->cfg.mode == SERBF_MODE_REFERENCE;
    s.dt = (T)dt;
    a.WL1 = sb.WL1;
    a.WL2 = sb.WL2;
    a.hloc = e->d_hloc;
    a.sbound = e->closure_kappa > 0.0 ? e->d_sbound : nullptr;
    a.ntr = e->ntr;
    a.hbar = (float)std::sqrt(8.0 * SERBF_PI / (std::sqrt(3.0) * e->N));
    a.umax = (float)(e->cfg.flow == SERBF_FLOW_DEFORM ? e->cfg.u0 + e->cfg.k_def : e->cfg.u0);
    if (!(a.umax > 0.f)) a.umax = 1.f;
    a.R = (float)e->cfg.R;
    a.dh = (float)(e->cfg.htop / e->cfg.Nv);
    return a;
}

// recomputed on every call (the stencils may have changed); allocated once per node set
int ensure_hloc(serbf_engine *e)
{
    if (!e->d_hloc) SERBF_HIP(hipMalloc((void **)&e->d_hloc, (size_t)e->N * sizeof(float)));
    if (!e->d_nmax) SERBF_HIP(hipMalloc((void **)&e->d_nmax, (size_t)e->N * 4 * sizeof(float)));
    hipLaunchKernelGGL(hloc_kernel, dim3((e->N + 255) / 256), dim3(256), 0, e->stream, e->N, e->n, e->d_x, e->d_y,
                       e->d_z, e->d_idx, e->d_hloc);
    SERBF_HIP(hipGetLastError());
    return SERBF_OK;
}

// row sums G_p = sum_j |L_pj| of the Delta^k weights
__global__ void rowsum_kernel(int N, int n, const double *WL, double *G)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    double s = 0.0;
    for (int k = 0; k < n; k++) s += fabs(WL[(size_t)p * n + k]);
    G[p] = s;
}

__global__ void bound_kernel(int N, const double *G, double num, float *b)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N) return;
    b[p] = G[p] > 0.0 ? (float)fmin(num / G[p], 65504.0) : 65504.f;
}

// b_p = kappa G_ref / G_p with G_ref the 99.9th percentile of the row sums (a node with a larger
// row sum than that gets a bound below kappa; the largest 0.1% of the rows do not set the scale)
int compute_bound(serbf_engine *e)
{
    SERBF_CHECK(e->have_weights && e->d_WL, SERBF_ERR_STATE, "closure bound: no Delta^k weights");
    const int N = e->N;
    if (!e->d_sbound) SERBF_HIP(hipMalloc((void **)&e->d_sbound, (size_t)N * sizeof(float)));
    DevTemp<double> G;
    SERBF_HIP(G.alloc(N));
    hipLaunchKernelGGL(rowsum_kernel, dim3((N + 255) / 256), dim3(256), 0, e->stream, N, e->n, e->d_WL, G.p);
    SERBF_HIP(hipGetLastError());
    std::vector<double> h(N);
    SERBF_HIP(hipMemcpyAsync(h.data(), G.p, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    const size_t k = (size_t)(0.999 * (N - 1));
    std::nth_element(h.begin(), h.begin() + k, h.end());
    const double gref = h[k];
    SERBF_CHECK(gref > 0.0, SERBF_ERR_STATE, "closure bound: Delta^k weights are zero");
    hipLaunchKernelGGL(bound_kernel, dim3((N + 255) / 256), dim3(256), 0, e->stream, N, G.p, e->closure_kappa * gref,
                       e->d_sbound);
    SERBF_HIP(hipGetLastError());
    SERBF_HIP(hipStreamSynchronize(e->stream));
    return SERBF_OK;
}

template <typename T, int S>
void launch_apply_s(serbf_engine *e, dim3 grid, size_t bytes, const FeatArgs<T> &a, const NetView &nv, T *gamma,
                    _Float16 *s16, T hv, float smax, int every, int npad)
{
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)apply_kernel<T, S>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr = true;
    }
    hipLaunchKernelGGL((apply_kernel<T, S>), grid, dim3(256), bytes, e->stream, a, nv, gamma, s16, hv, smax, every,
                       npad);
}

// the tile form of the evaluation: FP32, 4 tracers of 64 levels, config-mode flows, a tile plan
// (the engine builds one for the tile stage kernels) whose weight blocks fit the LDS
bool cvec_supported(const serbf_engine *e)
{
    return e->have_plan && mtile_supported(e) && e->ntr == 4 && vertical_window_fits(e) &&
           (size_t)e->N * 4 * 64 * 4 < (1ull << 32) &&
           cvec_lds_bytes((e->tile_umax + 15) & ~15) <= (size_t)160 * 1024 && e->d_nmax;
}

template <bool S16>
void launch_cvec_s(serbf_engine *e, const FeatArgs<float> &a, const NetView &nv, const MTilePlan &tp, int up,
                   size_t bytes, float *gamma, _Float16 *s16, float hv, float smax, int every)
{
    static bool attr = false;
    if (!attr) {
        (void)hipFuncSetAttribute((const void *)cvec_kernel<S16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr = true;
    }
    const size_t nb = ((size_t)e->N * 64 + 255) / 256;
    hipLaunchKernelGGL(nodemax_kernel, dim3((unsigned)std::min<size_t>(nb, 4096)), dim3(256), 0, e->stream, a, e->d_nmax,
                       every);
    hipLaunchKernelGGL((cvec_kernel<S16>), dim3(e->ntiles), dim3(256), bytes, e->stream, a, nv, tp, up,
                       (const float *)e->d_nmax, gamma, s16, hv, smax, every);
}

int launch_cvec(serbf_engine *e, const FeatArgs<float> &a0, float *gamma, int every, _Float16 *s16)
{
    FeatArgs<float> a = a0;
    a.s.nblk = e->ntiles;
    MTilePlan tp;
    tp.uni = e->d_tuni;
    tp.ucount = e->d_tcount;
    tp.lidx = e->d_lidx;
    tp.ustride = e->tile_ustride;
    const int up = (e->tile_umax + 15) & ~15;
    const size_t bytes = cvec_lds_bytes(up);
    const NetView nv = view(e->closure);
    if (s16) launch_cvec_s<true>(e, a, nv, tp, up, bytes, gamma, s16, (float)e->cfg.hv_coef, (float)e->closure_smax, every);
    else launch_cvec_s<false>(e, a, nv, tp, up, bytes, gamma, s16, (float)e->cfg.hv_coef, (float)e->closure_smax, every);
    SERBF_HIP(hipGetLastError());
    return SERBF_OK;
}

template <typename T>
// writes the coefficients to gamma, or (s16 != null) the multipliers in FP16 to s16
int launch_apply(serbf_engine *e, const FeatArgs<T> &a, T *gamma, int every, _Float16 *s16 = nullptr)
{
    if constexpr (std::is_same<T, float>::value)
        if (cvec_supported(e)) return launch_cvec(e, a, gamma, every, s16);
    const int npad = (e->n + 7) / 8 * 8;
    const int npw = 64 / e->S;
    const long passes = ((long)e->N + npw - 1) / npw * ((e->ntr + 3) / 4);
    // a fixed grid of 8 blocks (32 wavefronts) per CU striding over the passes
    if (e->cus <= 0) {
        hipDeviceProp_t prop;
        e->cus = hipGetDeviceProperties(&prop, e->device) == hipSuccess ? prop.multiProcessorCount : 256;
    }
    long blocks = (long)e->cus * 8;
    if (blocks > (passes + 3) / 4) blocks = (passes + 3) / 4;
    const dim3 grid((unsigned)blocks);
    const NetView nv = view(e->closure);
    const T hv = (T)e->cfg.hv_coef;
    const float smax = (float)e->closure_smax;
    if (e->S == 16) {
        const size_t bytes = 4 * ApplyGeom<T, 16>::wave_bytes(npad);
        SERBF_CHECK(bytes <= 160 * 1024, SERBF_ERR_UNSUPPORTED, "closure: stencil too large for the LDS");
        launch_apply_s<T, 16>(e, grid, bytes, a, nv, gamma, s16, hv, smax, every, npad);
    } else {
        const size_t bytes = 4 * ApplyGeom<T, 64>::wave_bytes(npad);
        SERBF_CHECK(bytes <= 160 * 1024, SERBF_ERR_UNSUPPORTED, "closure: stencil too large for the LDS");
        launch_apply_s<T, 64>(e, grid, bytes, a, nv, gamma, s16, hv, smax, every, npad);
    }
    SERBF_HIP(hipGetLastError());
    return SERBF_OK;
}

int alloc_gamma(serbf_engine *e)
{
    const size_t P = (size_t)e->N * e->ntr * e->S;
    if (!e->s64.gamma) {
        SERBF_HIP(hipMalloc((void **)&e->s64.gamma, P * sizeof(double)));
        SERBF_HIP(hipMemsetAsync(e->s64.gamma, 0, P * sizeof(double), e->stream));
    }
    if (!e->s32.gamma) {
        SERBF_HIP(hipMalloc((void **)&e->s32.gamma, P * sizeof(float)));
        SERBF_HIP(hipMemsetAsync(e->s32.gamma, 0, P * sizeof(float), e->stream));
    }
    if (!e->s32.s16) {
        SERBF_HIP(hipMalloc((void **)&e->s32.s16, P * sizeof(_Float16)));
        SERBF_HIP(hipMemsetAsync(e->s32.s16, 0, P * sizeof(_Float16), e->stream));
    }
    return SERBF_OK;
}

}  // namespace

template <typename T>
int closure_enqueue(serbf_engine *e, StateBuffers<T> &sb, double dt, int step_off, bool compact)
{
    FeatArgs<T> a = feat_args<T>(e, sb, sb.qT, dt, sb.tscal, e->d_base, step_off);
    a.origin = e->d_base + 2;
    return launch_apply<T>(e, a, sb.gamma, e->closure_every, compact ? sb.s16 : nullptr);
}
template int closure_enqueue<float>(serbf_engine *, StateBuffers<float> &, double, int, bool);
template int closure_enqueue<double>(serbf_engine *, StateBuffers<double> &, double, int, bool);

}  // namespace serbf

using namespace serbf;

extern "C" {

static int closure_init(serbf_closure *c)
{
    SERBF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    SERBF_HIP(hipMalloc((void **)&c->params, kP * sizeof(float)));
    SERBF_HIP(hipMalloc((void **)&c->adam_m, kP * sizeof(float)));
    SERBF_HIP(hipMalloc((void **)&c->adam_v, kP * sizeof(float)));
    SERBF_HIP(hipMalloc((void **)&c->w1b, kH * 16 * sizeof(unsigned short)));
    SERBF_HIP(hipMalloc((void **)&c->w2b, kH * kH * sizeof(unsigned short)));
    SERBF_HIP(hipMalloc((void **)&c->w2tb, kH * kH * sizeof(unsigned short)));
    SERBF_HIP(hipMalloc((void **)&c->norm, 32 * sizeof(float)));
    SERBF_HIP(hipMalloc((void **)&c->partials, (size_t)(kMaxWaves + 4) * kPS * sizeof(float)));
    SERBF_HIP(hipMalloc((void **)&c->grad, kPS * sizeof(float)));
    return SERBF_OK;
}

int serbf_closure_create(serbf_closure **out, int device)
{
    SERBF_CHECK(out, SERBF_ERR_ARG, "closure_create: null");
    *out = nullptr;
    SERBF_HIP(hipSetDevice(device));
    serbf_closure *c = new serbf_closure();
    c->device = device;
    const int rc = closure_init(c);
    if (rc) {
        serbf_closure_destroy(c);   // releases what was created before the failure
        return rc;
    }
    *out = c;
    return SERBF_OK;
}

void serbf_closure_destroy(serbf_closure *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (float *p : {c->params, c->adam_m, c->adam_v, c->norm, c->feat, c->A, c->R, c->scap, c->partials, c->grad,
                     c->losses, c->out})
        if (p) (void)hipFree(p);
    for (unsigned short *p : {c->w1b, c->w2b, c->w2tb})
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int serbf_closure_set_params(serbf_closure *c, const float *params, const float *norm)
{
    SERBF_CHECK(c && params && norm, SERBF_ERR_ARG, "closure_set_params: null");
    SERBF_HIP(hipSetDevice(c->device));
    SERBF_HIP(hipMemcpyAsync(c->params, params, kP * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SERBF_HIP(hipMemcpyAsync(c->norm, norm, 32 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SERBF_HIP(hipMemsetAsync(c->adam_m, 0, kP * sizeof(float), c->stream));
    SERBF_HIP(hipMemsetAsync(c->adam_v, 0, kP * sizeof(float), c->stream));
    hipLaunchKernelGGL(pack_kernel, dim3((kP + 255) / 256), dim3(256), 0, c->stream, c->params, c->w1b, c->w2b,
                       c->w2tb);
    SERBF_HIP(hipGetLastError());
    SERBF_HIP(hipStreamSynchronize(c->stream));
    c->step = 0;
    c->have_params = true;
    return SERBF_OK;
}

int serbf_closure_set_smax(serbf_closure *c, double s_max)
{
    SERBF_CHECK(c && s_max > 1.0, SERBF_ERR_ARG, "closure_set_smax: s_max must exceed 1");
    c->s_max = s_max;
    return SERBF_OK;
}

int serbf_closure_get_params(serbf_closure *c, float *params)
{
    SERBF_CHECK(c && params && c->have_params, SERBF_ERR_STATE, "closure_get_params: no parameters");
    SERBF_HIP(hipSetDevice(c->device));
    SERBF_HIP(hipMemcpyAsync(params, c->params, kP * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SERBF_HIP(hipStreamSynchronize(c->stream));
    return SERBF_OK;
}

int serbf_closure_set_data(serbf_closure *c, long n, const float *feat, const float *A, const float *R)
{
    SERBF_CHECK(c && n > 0 && feat && A && R, SERBF_ERR_ARG, "closure_set_data: bad arguments");
    SERBF_HIP(hipSetDevice(c->device));
    if (n > c->cap) {
        // a failed allocation leaves no samples and no capacity, not a stale size
        c->n = 0;
        c->cap = 0;
        for (float **p : {&c->feat, &c->A, &c->R, &c->scap})
            if (*p) { (void)hipFree(*p); *p = nullptr; }
        SERBF_HIP(hipMalloc((void **)&c->feat, (size_t)n * 16 * sizeof(float)));
        SERBF_HIP(hipMalloc((void **)&c->A, (size_t)n * sizeof(float)));
        SERBF_HIP(hipMalloc((void **)&c->R, (size_t)n * sizeof(float)));
        c->cap = n;
    }
    SERBF_HIP(hipMemcpyAsync(c->feat, feat, (size_t)n * 16 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SERBF_HIP(hipMemcpyAsync(c->A, A, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SERBF_HIP(hipMemcpyAsync(c->R, R, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SERBF_HIP(hipStreamSynchronize(c->stream));
    c->n = n;
    if (c->scap) { (void)hipFree(c->scap); c->scap = nullptr; }   // caps belong to the previous samples
    return SERBF_OK;
}

int serbf_closure_set_caps(serbf_closure *c, long n, const float *cap)
{
    SERBF_CHECK(c, SERBF_ERR_ARG, "closure_set_caps: null");
    SERBF_HIP(hipSetDevice(c->device));
    if (!cap) {
        if (c->scap) { (void)hipFree(c->scap); c->scap = nullptr; }
        return SERBF_OK;
    }
    SERBF_CHECK(n == c->n && n > 0, SERBF_ERR_ARG, "closure_set_caps: one cap per sample of the training set");
    // set_data drops the caps, so an existing buffer holds exactly c->n values
    if (!c->scap) SERBF_HIP(hipMalloc((void **)&c->scap, (size_t)n * sizeof(float)));
    SERBF_HIP(hipMemcpyAsync(c->scap, cap, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    SERBF_HIP(hipStreamSynchronize(c->stream));
    return SERBF_OK;
}

int serbf_closure_grad(serbf_closure *c, long first, long count, float *grad)
{
    SERBF_CHECK(c && grad, SERBF_ERR_ARG, "closure_grad: null");
    SERBF_HIP(hipSetDevice(c->device));
    int rc = run_grad(c, first, count, false, 0.f, -1);
    if (rc) return rc;
    SERBF_HIP(hipMemcpyAsync(grad, c->grad, (kP + 1) * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SERBF_HIP(hipStreamSynchronize(c->stream));
    return SERBF_OK;
}

int serbf_closure_train(serbf_closure *c, long first, long count, int steps, int batch, double lr, float *losses)
{
    SERBF_CHECK(c && steps >= 0 && batch > 0 && batch % 16 == 0, SERBF_ERR_ARG, "closure_train: bad arguments");
    SERBF_CHECK(first >= 0 && count >= batch && first + count <= c->n, SERBF_ERR_ARG,
                "closure_train: the sample range must lie in the set and hold at least one batch");
    SERBF_HIP(hipSetDevice(c->device));
    if (steps > c->loss_cap) {
        if (c->losses) { (void)hipFree(c->losses); c->losses = nullptr; }
        c->loss_cap = 0;
        SERBF_HIP(hipMalloc((void **)&c->losses, (size_t)steps * sizeof(float)));
        c->loss_cap = steps;
    }
    const long nb = count / batch;
    for (int s = 0; s < steps; s++) {
        const long b0 = first + (long)((c->step) % nb) * batch;
        int rc = run_grad(c, b0, batch, true, (float)lr, s);
        if (rc) return rc;
    }
    if (losses && steps > 0)
        SERBF_HIP(hipMemcpyAsync(losses, c->losses, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SERBF_HIP(hipStreamSynchronize(c->stream));
    return SERBF_OK;
}

int serbf_closure_eval(serbf_closure *c, long first, long count, float *o)
{
    SERBF_CHECK(c && o && c->have_params, SERBF_ERR_STATE, "closure_eval: no parameters");
    SERBF_CHECK(count > 0 && count % 16 == 0 && first >= 0 && first + count <= c->n, SERBF_ERR_ARG,
                "closure_eval: sample range must lie in the set and be a multiple of 16");
    SERBF_HIP(hipSetDevice(c->device));
    if (count > c->out_cap) {
        if (c->out) { (void)hipFree(c->out); c->out = nullptr; }
        c->out_cap = 0;
        SERBF_HIP(hipMalloc((void **)&c->out, (size_t)count * sizeof(float)));
        c->out_cap = count;
    }
    const int ntiles = (int)(count / 16);
    hipLaunchKernelGGL(eval_kernel, dim3(launch_grid_waves(ntiles)), dim3(256), 0, c->stream, view(c), c->feat,
                       first, ntiles, c->out);
    SERBF_HIP(hipGetLastError());
    SERBF_HIP(hipMemcpyAsync(o, c->out, (size_t)count * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SERBF_HIP(hipStreamSynchronize(c->stream));
    return SERBF_OK;
}

int serbf_closure_features(serbf_engine *e, double t, double dt, float *feat, double *lap)
{
    SERBF_CHECK(e && feat && lap, SERBF_ERR_ARG, "closure_features: null");
    SERBF_CHECK(e->have_state && e->have_lap && e->s64.WL1, SERBF_ERR_STATE,
                "closure_features: needs a state and config-mode weights with hv_order >= 3");
    SERBF_HIP(hipSetDevice(e->device));
    int rc = ensure_hloc(e);
    if (rc) return rc;
    double ts[4 * kTimeScalars] = {0};
    time_scalars_public(e->cfg, t, ts);
    const size_t nsm = (size_t)e->N * e->ntr * e->cfg.Nv;
    DevTemp<double> dts, dlap;
    DevTemp<float> dfeat;
    SERBF_HIP(dts.alloc(4 * kTimeScalars));
    SERBF_HIP(dlap.alloc(nsm));
    SERBF_HIP(dfeat.alloc(nsm * 16));
    double *d_ts = dts.p, *d_lap = dlap.p;
    float *d_feat = dfeat.p;
    SERBF_HIP(hipMemcpyAsync(d_ts, ts, sizeof(ts), hipMemcpyHostToDevice, e->stream));
    const FeatArgs<double> a = feat_args<double>(e, e->s64, e->s64.qT, dt, d_ts, e->d_base + 1, 0);
    const long ncol = (long)e->N * e->ntr;
    const long waves = e->S == 16 ? (ncol + 3) / 4 : ncol;
    const dim3 grid((unsigned)((waves + 3) / 4));
    if (e->S == 16) hipLaunchKernelGGL((features_kernel<double, 16>), grid, dim3(256), 0, e->stream, a, d_feat, d_lap);
    else hipLaunchKernelGGL((features_kernel<double, 64>), grid, dim3(256), 0, e->stream, a, d_feat, d_lap);
    SERBF_HIP(hipGetLastError());
    SERBF_HIP(hipMemcpyAsync(feat, d_feat, nsm * 16 * sizeof(float), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipMemcpyAsync(lap, d_lap, nsm * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    return SERBF_OK;
}

int serbf_closure_attach(serbf_engine *e, serbf_closure *c, int every, double s_max)
{
    SERBF_CHECK(e, SERBF_ERR_ARG, "closure_attach: null engine");
    SERBF_HIP(hipSetDevice(e->device));
    if (e->graph.exec) { (void)hipGraphExecDestroy(e->graph.exec); e->graph.exec = nullptr; }
    if (!c) {
        e->closure = nullptr;
        e->hv_mode = 0;
        return SERBF_OK;
    }
    // the FP32 tile kernel takes the multiplier in FP16, whose largest finite value is 65504
    SERBF_CHECK(every >= 1 && every <= 4096 && s_max > 0.0 && s_max <= 65504.0, SERBF_ERR_ARG,
                "closure_attach: every 1..4096, 0 < s_max <= 65504");
    SERBF_CHECK(c->have_params, SERBF_ERR_STATE, "closure_attach: closure has no parameters");
    SERBF_CHECK(c->device == e->device, SERBF_ERR_ARG, "closure_attach: closure and engine on different devices");
    SERBF_CHECK(e->have_state && e->have_lap && e->s64.WL1 && e->s32.WL1, SERBF_ERR_STATE,
                "closure_attach: needs a state and config-mode weights with hv_order >= 3");
    SERBF_CHECK(e->cfg.hv_order > 0, SERBF_ERR_STATE, "closure_attach: hyperviscosity order is 0");
    int rc = ensure_hloc(e);
    if (!rc) rc = alloc_gamma(e);
    if (rc) return rc;
    SERBF_HIP(hipStreamSynchronize(c->stream));
    e->closure = c;
    e->closure_start = e->steps_done;   // the first step after attaching evaluates
    e->closure_every = every;
    e->closure_smax = s_max;
    e->hv_mode = 1;
    return e->closure_kappa > 0.0 ? compute_bound(e) : SERBF_OK;
}

int serbf_closure_bound(serbf_engine *e, double kappa, float *bound)
{
    SERBF_CHECK(e && kappa >= 0.0 && std::isfinite(kappa), SERBF_ERR_ARG, "closure_bound: kappa >= 0");
    SERBF_HIP(hipSetDevice(e->device));
    if (e->graph.exec) { (void)hipGraphExecDestroy(e->graph.exec); e->graph.exec = nullptr; }
    e->closure_kappa = kappa;
    if (kappa == 0.0) return SERBF_OK;
    int rc = compute_bound(e);
    if (rc) {
        e->closure_kappa = 0.0;
        return rc;
    }
    if (bound) {
        SERBF_HIP(hipMemcpyAsync(bound, e->d_sbound, (size_t)e->N * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        SERBF_HIP(hipStreamSynchronize(e->stream));
    }
    return SERBF_OK;
}

int serbf_closure_gamma(serbf_engine *e, int precision, double t, double dt, double *gamma)
{
    SERBF_CHECK(e && gamma && e->closure, SERBF_ERR_STATE, "closure_gamma: no closure attached");
    SERBF_HIP(hipSetDevice(e->device));
    double ts[4 * kTimeScalars] = {0};
    time_scalars_public(e->cfg, t, ts);
    const int S = e->S, Nv = e->cfg.Nv;
    const size_t cols = (size_t)e->N * e->ntr, P = cols * S;
    int rc = SERBF_OK;
    if (precision == SERBF_F32) {
        // the time loop's FP32 evaluation runs on the tile plan where one applies: build it here too
        if (!e->have_plan && e->ntr == 4 && e->S == 64 && e->n <= kMaxStencil && e->cfg.mode == SERBF_MODE_CONFIG &&
            (rc = tile_plan(e)))
            return rc;
        float tsf[4 * kTimeScalars];
        for (int i = 0; i < 4 * kTimeScalars; i++) tsf[i] = (float)ts[i];
        DevTemp<float> d_ts;
        SERBF_HIP(d_ts.alloc(4 * kTimeScalars));
        SERBF_HIP(hipMemcpyAsync(d_ts.p, tsf, sizeof(tsf), hipMemcpyHostToDevice, e->stream));
        const FeatArgs<float> a = feat_args<float>(e, e->s32, e->s32.qT, dt, d_ts.p, e->d_base + 1, 0);
        rc = launch_apply<float>(e, a, e->s32.gamma, 1);
        if (!rc) {
            std::vector<float> g(P);
            SERBF_HIP(hipMemcpyAsync(g.data(), e->s32.gamma, P * sizeof(float), hipMemcpyDeviceToHost, e->stream));
            SERBF_HIP(hipStreamSynchronize(e->stream));
            for (size_t col = 0; col < cols; col++)
                for (int v = 0; v < Nv; v++) gamma[col * Nv + v] = g[col * S + v];
        }
    } else {
        DevTemp<double> d_ts;
        SERBF_HIP(d_ts.alloc(4 * kTimeScalars));
        SERBF_HIP(hipMemcpyAsync(d_ts.p, ts, sizeof(ts), hipMemcpyHostToDevice, e->stream));
        const FeatArgs<double> a = feat_args<double>(e, e->s64, e->s64.qT, dt, d_ts.p, e->d_base + 1, 0);
        rc = launch_apply<double>(e, a, e->s64.gamma, 1);
        if (!rc) {
            rc = download_padded(e, gamma, e->s64.gamma, cols, Nv, S);
        }
    }
    return rc;
}

}  // extern "C"
