// Engine API of libserbf: device-resident node set, stencils, weights and state.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"

namespace serbf {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int ensure_scratch(serbf_engine *e, size_t bytes)
{
    if (bytes <= e->scratch_bytes) return SERBF_OK;
    if (e->d_scratch) (void)hipFree(e->d_scratch);
    e->d_scratch = nullptr;
    e->scratch_bytes = 0;
    SERBF_HIP(hipMalloc(&e->d_scratch, bytes));
    e->scratch_bytes = bytes;
    return SERBF_OK;
}

// Block partial sums of the verification norms (wrappers.c:133-157):
// {sum|d|, sum|qr|, sum d^2, sum qr^2, mass_ref, mass, max|d|, max|q|, max q, min q}
template <typename T>
__global__ void norms_kernel(size_t ncols, int Nv, int S, const double *qr, const T *q, const double *wmass,
                             double *part)
{
    __shared__ double red[10][256];
    double a[10] = {0, 0, 0, 0, 0, 0, 0, 0, -1e300, 1e300};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ncols * S; i += (size_t)gridDim.x * blockDim.x) {
        const int v = (int)(i % S);
        if (v >= Nv) continue;
        const double r = qr[i], x = (double)q[i], d = fabs(r - x);
        a[0] += d; a[1] += fabs(r); a[2] += d * d; a[3] += r * r;
        a[4] += wmass[v] * r; a[5] += wmass[v] * x;
        a[6] = fmax(a[6], d); a[7] = fmax(a[7], fabs(x)); a[8] = fmax(a[8], x); a[9] = fmin(a[9], x);
    }
    for (int k = 0; k < 10; k++) red[k][threadIdx.x] = a[k];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 6; k++) red[k][threadIdx.x] += red[k][threadIdx.x + s];
            for (int k = 6; k < 9; k++) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + s]);
            red[9][threadIdx.x] = fmin(red[9][threadIdx.x], red[9][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 10; k++) part[blockIdx.x * 10 + k] = red[k][0];
}

// Plain streaming copy (16 B per lane per iteration): the attainable-bandwidth and
// HBM-counter calibration reference.
typedef float copy_v4 __attribute__((ext_vector_type(4)));

// dst = src in 16-byte lanes; each thread moves U vectors per pass (all U loads issued before
// the first store), with non-temporal loads and stores when NT
template <int U, bool NT>
__global__ void __launch_bounds__(256) stream_copy_kernel(size_t n16, const copy_v4 *src, copy_v4 *dst)
{
    const size_t step = (size_t)gridDim.x * 256 * U;
    for (size_t base = (size_t)blockIdx.x * 256 * U + threadIdx.x; base < n16; base += step) {
        copy_v4 v[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const size_t i = base + (size_t)k * 256;
            if (i < n16) v[k] = NT ? __builtin_nontemporal_load(src + i) : src[i];
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
            const size_t i = base + (size_t)k * 256;
            if (i < n16) {
                if (NT) __builtin_nontemporal_store(v[k], dst + i);
                else dst[i] = v[k];
            }
        }
    }
}

template <int U, bool NT>
void launch_copy(hipStream_t s, int blocks, size_t n16, const copy_v4 *a, copy_v4 *b)
{
    hipLaunchKernelGGL((stream_copy_kernel<U, NT>), dim3(blocks), dim3(256), 0, s, n16, a, b);
}

}  // namespace serbf

using namespace serbf;

namespace serbf {

int upload_padded(serbf_engine *e, double *dst, const double *src, size_t cols, int Nv, int S)
{
    if (Nv == S) {
        SERBF_HIP(hipMemcpyAsync(dst, src, cols * S * sizeof(double), hipMemcpyHostToDevice, e->stream));
        SERBF_HIP(hipStreamSynchronize(e->stream));
        return SERBF_OK;
    }
    std::vector<double> tmp(cols * S, 0.0);
    for (size_t c = 0; c < cols; c++) std::memcpy(&tmp[c * S], src + c * Nv, Nv * sizeof(double));
    SERBF_HIP(hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    return SERBF_OK;
}

int download_padded(serbf_engine *e, double *dst, const double *src, size_t cols, int Nv, int S)
{
    if (Nv == S) {
        SERBF_HIP(hipMemcpyAsync(dst, src, cols * S * sizeof(double), hipMemcpyDeviceToHost, e->stream));
        SERBF_HIP(hipStreamSynchronize(e->stream));
        return SERBF_OK;
    }
    std::vector<double> tmp(cols * S);
    SERBF_HIP(hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    for (size_t c = 0; c < cols; c++) std::memcpy(dst + c * Nv, &tmp[c * S], Nv * sizeof(double));
    return SERBF_OK;
}

}  // namespace serbf

extern "C" {

const char *serbf_last_error(void) { return g_err.c_str(); }

static int engine_init(serbf_engine *e)
{
    SERBF_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    SERBF_HIP(hipEventCreate(&e->ev0));
    SERBF_HIP(hipEventCreate(&e->ev1));
    SERBF_HIP(hipMalloc((void **)&e->d_base, 3 * sizeof(int)));
    SERBF_HIP(hipMemset(e->d_base, 0, 3 * sizeof(int)));
    return SERBF_OK;
}

int serbf_create(serbf_engine **out, int device)
{
    SERBF_CHECK(out, SERBF_ERR_ARG, "create: null out");
    *out = nullptr;
    int ndev = 0;
    SERBF_HIP(hipGetDeviceCount(&ndev));
    SERBF_CHECK(device >= 0 && device < ndev, SERBF_ERR_ARG, "create: no such device");
    SERBF_HIP(hipSetDevice(device));
    serbf_engine *e = new serbf_engine();
    e->device = device;
    const int rc = engine_init(e);
    if (rc) {
        serbf_destroy(e);   // releases what was created before the failure
        return rc;
    }
    *out = e;
    return SERBF_OK;
}

static void clear_sizes(serbf_engine *e);

static void free_nodes(serbf_engine *e)
{
    for (double *p : {e->d_x, e->d_y, e->d_z, e->d_lam, e->d_phi, e->d_D, e->d_Wx, e->d_Wy, e->d_WL, e->d_WL1,
                      e->d_WL2, e->d_q0, e->d_wmass})
        if (p) (void)hipFree(p);
    for (int *p : {e->d_idx, e->d_status, e->d_vstart, e->d_tuni, e->d_tcount})
        if (p) (void)hipFree(p);
    if (e->d_lidx) (void)hipFree(e->d_lidx);
    clear_sizes(e);
    if (e->d_hloc) (void)hipFree(e->d_hloc);
    e->d_hloc = nullptr;
    if (e->d_nmax) (void)hipFree(e->d_nmax);
    e->d_nmax = nullptr;
    if (e->d_sbound) (void)hipFree(e->d_sbound);
    e->d_sbound = nullptr;
    e->closure_kappa = 0.0;
    e->closure = nullptr;
    e->hv_mode = 0;
    e->d_tuni = e->d_tcount = nullptr;
    e->d_lidx = nullptr;
    e->have_plan = false;
    e->d_x = e->d_y = e->d_z = e->d_lam = e->d_phi = e->d_D = e->d_Wx = e->d_Wy = e->d_WL = e->d_WL1 = e->d_WL2 =
        e->d_q0 = e->d_wmass = nullptr;
    e->have_lap = false;
    e->d_idx = e->d_status = e->d_vstart = nullptr;
    rk4_free(e);
    e->have_stencils = e->have_weights = e->have_state = e->f32_ready = false;
}

void serbf_destroy(serbf_engine *e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    free_nodes(e);
    if (e->d_base) (void)hipFree(e->d_base);
    if (e->d_scratch) (void)hipFree(e->d_scratch);
    if (e->ev0) (void)hipEventDestroy(e->ev0);
    if (e->ev1) (void)hipEventDestroy(e->ev1);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int serbf_set_nodes(serbf_engine *e, int N, const double *x, const double *y, const double *z)
{
    SERBF_CHECK(e && N > 0 && x && y && z, SERBF_ERR_ARG, "set_nodes: bad arguments");
    SERBF_HIP(hipSetDevice(e->device));
    free_nodes(e);
    e->configured = false;
    e->N = 0;   // no node set until every copy below has landed
    for (double **p : {&e->d_x, &e->d_y, &e->d_z}) SERBF_HIP(hipMalloc((void **)p, N * sizeof(double)));
    SERBF_HIP(hipMemcpyAsync(e->d_x, x, N * sizeof(double), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipMemcpyAsync(e->d_y, y, N * sizeof(double), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipMemcpyAsync(e->d_z, z, N * sizeof(double), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    e->N = N;
    return SERBF_OK;
}

// Optional longitude/latitude per node (the reference reads them from the node file);
// without them the velocity basis derives them from the coordinates.
int serbf_set_lonlat(serbf_engine *e, const double *lam, const double *phi)
{
    SERBF_CHECK(e && e->N > 0 && lam && phi, SERBF_ERR_ARG, "set_lonlat: bad arguments");
    SERBF_HIP(hipSetDevice(e->device));
    for (double **p : {&e->d_lam, &e->d_phi})
        if (!*p) SERBF_HIP(hipMalloc((void **)p, e->N * sizeof(double)));
    SERBF_HIP(hipMemcpyAsync(e->d_lam, lam, e->N * sizeof(double), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipMemcpyAsync(e->d_phi, phi, e->N * sizeof(double), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    return SERBF_OK;
}

int serbf_knn(serbf_engine *e, int n)
{
    SERBF_CHECK(e && e->N > 0, SERBF_ERR_STATE, "knn: no nodes");
    SERBF_CHECK(n >= 1 && n <= e->N && n <= kMaxStencil, SERBF_ERR_ARG, "knn: n out of range");
    SERBF_HIP(hipSetDevice(e->device));
    return knn_run(e, n);
}

static void clear_sizes(serbf_engine *e)
{
    for (int *p : {e->d_nn, e->d_deg, e->d_worder})
        if (p) (void)hipFree(p);
    e->d_nn = e->d_deg = e->d_worder = nullptr;
    e->h_nn.clear();
    e->h_deg.clear();
}

static int alloc_stencils_impl(serbf_engine *e, int n)
{
    e->have_plan = false;
    clear_sizes(e);
    if (e->n != n) {
        if (e->d_idx) (void)hipFree(e->d_idx);
        if (e->d_D) (void)hipFree(e->d_D);
        e->d_idx = nullptr;
        e->d_D = nullptr;
    }
    e->n = n;
    // each buffer on its own, so a call after a failed allocation completes the pair
    if (!e->d_idx) SERBF_HIP(hipMalloc((void **)&e->d_idx, (size_t)e->N * n * sizeof(int)));
    if (!e->d_D) SERBF_HIP(hipMalloc((void **)&e->d_D, (size_t)e->N * n * sizeof(double)));
    return SERBF_OK;
}

int serbf_get_stencils(serbf_engine *e, int *idx, double *D)
{
    SERBF_CHECK(e && e->have_stencils, SERBF_ERR_STATE, "get_stencils: none");
    const size_t m = (size_t)e->N * e->n;
    if (idx) SERBF_HIP(hipMemcpyAsync(idx, e->d_idx, m * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    if (D) SERBF_HIP(hipMemcpyAsync(D, e->d_D, m * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    return SERBF_OK;
}

int serbf_set_stencils(serbf_engine *e, int n, const int *idx)
{
    SERBF_CHECK(e && e->N > 0 && idx && n >= 1 && n <= kMaxStencil, SERBF_ERR_ARG, "set_stencils: bad arguments");
    for (size_t i = 0; i < (size_t)e->N * n; i++)
        SERBF_CHECK(idx[i] >= 0 && idx[i] < e->N, SERBF_ERR_ARG, "set_stencils: index out of range");
    for (int i = 0; i < e->N; i++)
        SERBF_CHECK(idx[(size_t)i * n] == i, SERBF_ERR_ARG, "set_stencils: stencils must start with the node itself");
    int rc = alloc_stencils_impl(e, n);
    if (rc) return rc;
    SERBF_HIP(hipMemcpyAsync(e->d_idx, idx, (size_t)e->N * n * sizeof(int), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    e->have_stencils = true;
    e->have_weights = false;
    return SERBF_OK;
}

int serbf_configure(serbf_engine *e, const serbf_config *cfg)
{
    SERBF_CHECK(e && cfg, SERBF_ERR_ARG, "configure: null");
    const serbf_config &c = *cfg;
    SERBF_CHECK(c.Nv >= 1 && c.Nv <= kMaxLevels, SERBF_ERR_UNSUPPORTED, "configure: Nv must be in [1, 64]");
    SERBF_CHECK(c.htop > 0 && c.R > 0, SERBF_ERR_ARG, "configure: htop, R must be positive");
    SERBF_CHECK(c.mode == SERBF_MODE_REFERENCE || c.mode == SERBF_MODE_CONFIG, SERBF_ERR_ARG, "configure: mode");
    SERBF_CHECK(c.flow >= SERBF_FLOW_TC2 && c.flow <= SERBF_FLOW_DEFORM, SERBF_ERR_ARG, "configure: flow");
    SERBF_CHECK(c.n >= 2 && c.n <= kMaxStencil, SERBF_ERR_ARG, "configure: n");
    const int ntr = c.ntracers > 0 ? c.ntracers : 1;
    SERBF_CHECK(ntr == 1 || ntr == 4, SERBF_ERR_UNSUPPORTED, "configure: 1 or 4 tracers");
    if (c.mode == SERBF_MODE_CONFIG) {
        SERBF_CHECK(c.nv >= 1 && c.nv <= kMaxNv && c.nv <= c.Nv, SERBF_ERR_ARG, "configure: nv");
        SERBF_CHECK(c.poly_deg >= 0 && c.poly_deg <= 6, SERBF_ERR_UNSUPPORTED, "configure: poly_deg");
        SERBF_CHECK(c.hv_order == 0 || c.poly_deg < 2 * c.hv_order, SERBF_ERR_UNSUPPORTED,
                    "configure: poly_deg must be below 2*hv_order");
        SERBF_CHECK(c.hv_order == 0 || (c.hv_phs_order == 0 && c.hv_eps > 0.0) ||
                        (c.hv_phs_order >= 2 * c.hv_order + 1 && (c.hv_phs_order & 1)),
                    SERBF_ERR_ARG,
                    "configure: hv_phs_order must be 0 (Gaussian, hv_eps > 0) or odd and >= 2k+1");
        SERBF_CHECK(c.phs_order & 1, SERBF_ERR_ARG, "configure: phs_order must be odd");
    } else {
        SERBF_CHECK(c.hv_order >= 0 && c.hv_order < SERBF_REF_MAX_HV_ORDER, SERBF_ERR_ARG, "configure: hv_order");
    }
    SERBF_HIP(hipSetDevice(e->device));
    rk4_free(e);
    // a configuration that fails below leaves the engine unconfigured, not half-updated
    e->configured = false;
    e->have_weights = false;
    e->have_state = false;
    e->cfg = c;
    e->ntr = ntr;
    e->S = c.Nv <= 16 ? 16 : 64;
    const int nv = vertical_table(c, e->h_vstart, e->h_vW);
    SERBF_CHECK(nv > 0, SERBF_ERR_SINGULAR, "configure: vertical stencil system is singular");
    e->nv = nv;
    std::vector<int> vs(e->S, 0);
    for (int v = 0; v < c.Nv; v++) vs[v] = e->h_vstart[v];
    if (!e->d_vstart) SERBF_HIP(hipMalloc((void **)&e->d_vstart, kMaxLevels * sizeof(int)));
    SERBF_HIP(hipMemcpyAsync(e->d_vstart, vs.data(), e->S * sizeof(int), hipMemcpyHostToDevice, e->stream));
    // per-level mass weights rho(h) 4 pi r^2 dh / N (wrappers.c:146-147)
    if (!e->d_wmass) SERBF_HIP(hipMalloc((void **)&e->d_wmass, kMaxLevels * sizeof(double)));
    std::vector<double> wm(kMaxLevels, 0.0);
    const double dh = c.htop / c.Nv, rho0 = SERBF_P0 / (SERBF_RD * SERBF_T0);
    for (int v = 0; v < c.Nv; v++) {
        const double h = dh * (v + .5), r = c.R + h;
        wm[v] = rho0 * std::exp(-(h * SERBF_GRAVITY) / (SERBF_RD * SERBF_T0)) * (4 * SERBF_PI * r * r * dh) /
                (e->N > 0 ? e->N : 1);
    }
    SERBF_HIP(hipMemcpyAsync(e->d_wmass, wm.data(), kMaxLevels * sizeof(double), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    e->configured = true;
    return SERBF_OK;
}

int serbf_weights(serbf_engine *e)
{
    SERBF_CHECK(e && e->configured && e->have_stencils, SERBF_ERR_STATE, "weights: configure and stencils first");
    SERBF_CHECK(e->cfg.n == e->n, SERBF_ERR_ARG, "weights: config n differs from the stencil size");
    SERBF_HIP(hipSetDevice(e->device));
    return weights_run(e);
}

int serbf_get_weights(serbf_engine *e, double *Dx, double *Dy, double *L, int *status)
{
    SERBF_CHECK(e && e->have_weights, SERBF_ERR_STATE, "get_weights: none");
    const size_t m = (size_t)e->N * e->n;
    if (Dx) SERBF_HIP(hipMemcpyAsync(Dx, e->d_Wx, m * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (Dy) SERBF_HIP(hipMemcpyAsync(Dy, e->d_Wy, m * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (L) SERBF_HIP(hipMemcpyAsync(L, e->d_WL, m * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (status && e->d_status)
        SERBF_HIP(hipMemcpyAsync(status, e->d_status, e->N * sizeof(int), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    return SERBF_OK;
}

}  // extern "C"

int serbf::alloc_weights(serbf_engine *e)
{
    const size_t m = (size_t)e->N * e->n;
    for (double **p : {&e->d_Wx, &e->d_Wy, &e->d_WL, &e->d_WL1, &e->d_WL2}) {
        if (*p) { (void)hipFree(*p); *p = nullptr; }
        SERBF_HIP(hipMalloc((void **)p, m * sizeof(double)));
    }
    if (!e->d_status) SERBF_HIP(hipMalloc((void **)&e->d_status, e->N * sizeof(int)));
    return SERBF_OK;
}

int serbf::alloc_stencils(serbf_engine *e, int n) { return alloc_stencils_impl(e, n); }

extern "C" {

int serbf_set_weights(serbf_engine *e, const double *Dx, const double *Dy, const double *L)
{
    SERBF_CHECK(e && e->configured && e->have_stencils && Dx && Dy, SERBF_ERR_STATE, "set_weights: bad state");
    SERBF_HIP(hipSetDevice(e->device));
    rk4_free(e);
    e->have_state = false;
    int rc = alloc_weights(e);
    if (rc) return rc;
    const size_t m = (size_t)e->N * e->n;
    SERBF_HIP(hipMemcpyAsync(e->d_Wx, Dx, m * sizeof(double), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipMemcpyAsync(e->d_Wy, Dy, m * sizeof(double), hipMemcpyHostToDevice, e->stream));
    if (L) SERBF_HIP(hipMemcpyAsync(e->d_WL, L, m * sizeof(double), hipMemcpyHostToDevice, e->stream));
    else SERBF_HIP(hipMemsetAsync(e->d_WL, 0, m * sizeof(double), e->stream));
    SERBF_HIP(hipMemsetAsync(e->d_status, 0, e->N * sizeof(int), e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    e->have_weights = true;
    e->have_lap = false;
    return SERBF_OK;
}

int serbf_set_stencil_sizes(serbf_engine *e, const int *nn, const int *deg)
{
    SERBF_CHECK(e && e->have_stencils && e->configured, SERBF_ERR_STATE,
                "set_stencil_sizes: stencils and configuration first");
    SERBF_HIP(hipSetDevice(e->device));
    clear_sizes(e);
    e->have_weights = false;
    rk4_free(e);
    e->have_state = false;
    if (!nn && !deg) return SERBF_OK;
    SERBF_CHECK(nn && deg, SERBF_ERR_ARG, "set_stencil_sizes: both arrays or neither");
    SERBF_CHECK(e->cfg.mode == SERBF_MODE_CONFIG, SERBF_ERR_UNSUPPORTED, "set_stencil_sizes: config mode only");
    for (int i = 0; i < e->N; i++) {
        SERBF_CHECK(nn[i] >= 2 && nn[i] <= e->n, SERBF_ERR_ARG, "set_stencil_sizes: nn[i] must be in [2, n]");
        SERBF_CHECK(deg[i] >= 0 && deg[i] <= 6 && (e->cfg.hv_order == 0 || deg[i] < 2 * e->cfg.hv_order),
                    SERBF_ERR_ARG, "set_stencil_sizes: deg[i] must be in [0, 6] and below 2*hv_order");
        SERBF_CHECK(nn[i] + (deg[i] + 1) * (deg[i] + 2) / 2 <= 112, SERBF_ERR_UNSUPPORTED,
                    "set_stencil_sizes: nn + polynomial terms must be <= 112");
    }
    e->h_nn.assign(nn, nn + e->N);
    e->h_deg.assign(deg, deg + e->N);
    SERBF_HIP(hipMalloc((void **)&e->d_nn, e->N * sizeof(int)));
    SERBF_HIP(hipMalloc((void **)&e->d_deg, e->N * sizeof(int)));
    SERBF_HIP(hipMemcpyAsync(e->d_nn, nn, e->N * sizeof(int), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipMemcpyAsync(e->d_deg, deg, e->N * sizeof(int), hipMemcpyHostToDevice, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    return SERBF_OK;
}

int serbf_weights_autotune(serbf_engine *e, int sample, double *report, int *nvariants)
{
    SERBF_CHECK(e && e->configured && e->have_stencils, SERBF_ERR_STATE, "weights_autotune: configure and stencils first");
    SERBF_CHECK(e->cfg.n == e->n, SERBF_ERR_ARG, "weights_autotune: config n differs from the stencil size");
    SERBF_HIP(hipSetDevice(e->device));
    e->have_weights = false;
    rk4_free(e);
    e->have_state = false;
    return weights_autotune(e, sample, report, nvariants);
}

int serbf_weights_set_variant(serbf_engine *e, int variant)
{
    SERBF_CHECK(e && variant >= 0 && variant < 6, SERBF_ERR_ARG, "weights_set_variant: 0 to 5");
    e->wvariant = variant;
    return SERBF_OK;
}

int serbf_weights_stats(serbf_engine *e, double *ms, double *systems, int *variant)
{
    SERBF_CHECK(e && e->wstats_variant >= 0, SERBF_ERR_STATE, "weights_stats: no weight solve yet");
    if (ms) *ms = e->wstats_ms;
    if (systems) *systems = e->wstats_systems;
    if (variant) *variant = e->wstats_variant;
    return SERBF_OK;
}

int serbf_get_lap_weights(serbf_engine *e, double *L1, double *L2)
{
    SERBF_CHECK(e && e->have_weights && e->have_lap, SERBF_ERR_STATE,
                "get_lap_weights: needs config-mode weights with hv_order >= 3 from serbf_weights");
    const size_t m = (size_t)e->N * e->n;
    if (L1) SERBF_HIP(hipMemcpyAsync(L1, e->d_WL1, m * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    if (L2) SERBF_HIP(hipMemcpyAsync(L2, e->d_WL2, m * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    return SERBF_OK;
}

int serbf_get_vertical(serbf_engine *e, int *start, double *W, int *nv)
{
    SERBF_CHECK(e && e->configured, SERBF_ERR_STATE, "get_vertical: not configured");
    if (start) std::memcpy(start, e->h_vstart, e->cfg.Nv * sizeof(int));
    if (W) std::memcpy(W, e->h_vW, (size_t)e->cfg.Nv * e->nv * sizeof(double));
    if (nv) *nv = e->nv;
    return SERBF_OK;
}

int serbf_set_state(serbf_engine *e, const double *q)
{
    SERBF_CHECK(e && q, SERBF_ERR_ARG, "set_state: null");
    SERBF_HIP(hipSetDevice(e->device));
    if (!e->s64.qT) {
        int rc = rk4_prepare(e);
        if (rc) return rc;
    }
    const int Nv = e->cfg.Nv, S = e->S;
    const size_t cols = (size_t)e->N * e->ntr;
    {
        const int rc = upload_padded(e, e->s64.qT, q, cols, Nv, S);
        if (rc) return rc;
    }
    if (!e->d_q0) {
        SERBF_HIP(hipMalloc((void **)&e->d_q0, cols * S * sizeof(double)));
        SERBF_HIP(hipMemsetAsync(e->d_q0, 0, cols * S * sizeof(double), e->stream));
    }
    SERBF_HIP(hipMemcpyAsync(e->d_q0, e->s64.qT, cols * S * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    e->have_state = true;
    e->steps_done = 0;
    e->closure_start = 0;
    return serbf_sync_f32_from_f64(e);
}

int serbf_sync_f32_from_f64(serbf_engine *e)
{
    SERBF_CHECK(e && e->have_state && e->f32_ready, SERBF_ERR_STATE, "sync_f32: no state");
    const size_t P = (size_t)e->N * e->ntr * e->S;
    int rc = convert_to_f32(e, P, e->s64.qT, e->s32.qT);
    if (rc) return rc;
    SERBF_HIP(hipStreamSynchronize(e->stream));
    return SERBF_OK;
}

int serbf_get_state(serbf_engine *e, int precision, double *q)
{
    SERBF_CHECK(e && e->have_state && q, SERBF_ERR_STATE, "get_state: no state");
    const int Nv = e->cfg.Nv, S = e->S;
    const size_t cols = (size_t)e->N * e->ntr;
    if (precision == SERBF_F32) {
        std::vector<float> tmp(cols * S);
        SERBF_HIP(hipMemcpyAsync(tmp.data(), e->s32.qT, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, e->stream));
        SERBF_HIP(hipStreamSynchronize(e->stream));
        for (size_t c = 0; c < cols; c++)
            for (int v = 0; v < Nv; v++) q[c * Nv + v] = tmp[c * S + v];
        return SERBF_OK;
    }
    return download_padded(e, q, e->s64.qT, cols, Nv, S);
}

int serbf_rhs(serbf_engine *e, double t, double *F)
{
    SERBF_CHECK(e && F, SERBF_ERR_ARG, "rhs: null");
    SERBF_HIP(hipSetDevice(e->device));
    return rk4_rhs(e, t, F);
}

int serbf_run(serbf_engine *e, int precision, int nsteps, double dt, double t0)
{
    SERBF_CHECK(e, SERBF_ERR_ARG, "run: null");
    SERBF_HIP(hipSetDevice(e->device));
    return rk4_run(e, precision, nsteps, dt, t0, false, nullptr, nullptr);
}

int serbf_run_timed(serbf_engine *e, int precision, int nsteps, double dt, double t0, double *ms, double *stage_ms)
{
    SERBF_CHECK(e && ms && stage_ms, SERBF_ERR_ARG, "run_timed: null");
    SERBF_HIP(hipSetDevice(e->device));
    return rk4_run(e, precision, nsteps, dt, t0, true, ms, stage_ms);
}

int serbf_stream_copy(serbf_engine *e, size_t bytes, int reps, double *ms_per_copy)
{
    SERBF_CHECK(e && bytes >= 16 && reps >= 1 && ms_per_copy, SERBF_ERR_ARG, "stream_copy: bad arguments");
    SERBF_HIP(hipSetDevice(e->device));
    const size_t n16 = bytes / 16;
    DevTemp<copy_v4> da, db;
    SERBF_HIP(da.alloc(n16));
    SERBF_HIP(db.alloc(n16));
    copy_v4 *a = da.p, *b = db.p;
    SERBF_HIP(hipMemsetAsync(a, 0, n16 * 16, e->stream));
    // the tuned copy: the fastest of vectors per thread {1, 4, 8} x plain / non-temporal x
    // {1024, 4096, 16384} work-groups
    typedef void (*Launch)(hipStream_t, int, size_t, const copy_v4 *, copy_v4 *);
    const Launch variants[] = {launch_copy<1, false>, launch_copy<4, false>, launch_copy<8, false>,
                               launch_copy<1, true>,  launch_copy<4, true>,  launch_copy<8, true>};
    const int grids[] = {1024, 4096, 16384};
    float best = 0.f;
    for (Launch L : variants)
        for (int blocks : grids) {
            L(e->stream, blocks, n16, a, b);
            SERBF_HIP(hipGetLastError());
            SERBF_HIP(hipEventRecord(e->ev0, e->stream));
            for (int r = 0; r < reps; r++) L(e->stream, blocks, n16, a, b);
            SERBF_HIP(hipEventRecord(e->ev1, e->stream));
            SERBF_HIP(hipStreamSynchronize(e->stream));
            float t = 0.f;
            SERBF_HIP(hipEventElapsedTime(&t, e->ev0, e->ev1));
            if (best == 0.f || t < best) best = t;
        }
    *ms_per_copy = best / reps;
    return SERBF_OK;
}

int serbf_set_hv_split(serbf_engine *e, int split)
{
    SERBF_CHECK(e, SERBF_ERR_ARG, "set_hv_split: null");
    e->hv_split = split ? 1 : 0;
    return SERBF_OK;
}

int serbf_set_kernel_version(serbf_engine *e, int version)
{
    SERBF_CHECK(e && (version == 2 || version == 5 || version == 7), SERBF_ERR_ARG, "set_kernel_version: 2, 5 or 7");
    e->kver = version;
    return SERBF_OK;
}

int serbf_tile_info(serbf_engine *e, int *ntiles, int *umax)
{
    SERBF_CHECK(e, SERBF_ERR_ARG, "tile_info: null");
    if (ntiles) *ntiles = e->have_plan ? e->ntiles : 0;
    if (umax) *umax = e->have_plan ? e->tile_umax : 0;
    return SERBF_OK;
}

int serbf_set_stage_unroll(serbf_engine *e, int unroll)
{
    SERBF_CHECK(e && (unroll == 4 || unroll == 8 || unroll == 16), SERBF_ERR_ARG, "set_stage_unroll: 4, 8 or 16");
    e->unroll = unroll;
    return SERBF_OK;
}

int serbf_set_gamma(serbf_engine *e, const double *gamma)
{
    SERBF_CHECK(e && e->have_state, SERBF_ERR_STATE, "set_gamma: state first");
    if (!gamma) {
        e->hv_mode = 0;
        return SERBF_OK;
    }
    const int Nv = e->cfg.Nv, S = e->S;
    const size_t cols = (size_t)e->N * e->ntr, P = cols * S;
    if (!e->s64.gamma) {
        SERBF_HIP(hipMalloc((void **)&e->s64.gamma, P * sizeof(double)));
        SERBF_HIP(hipMalloc((void **)&e->s32.gamma, P * sizeof(float)));
        SERBF_HIP(hipMemsetAsync(e->s64.gamma, 0, P * sizeof(double), e->stream));
    }
    int rc = upload_padded(e, e->s64.gamma, gamma, cols, Nv, S);
    if (rc) return rc;
    rc = convert_to_f32(e, P, e->s64.gamma, e->s32.gamma);
    if (rc) return rc;
    SERBF_HIP(hipStreamSynchronize(e->stream));
    e->hv_mode = 1;
    return SERBF_OK;
}

int serbf_norms(serbf_engine *e, int precision, const double *q_ref, double *out)
{
    SERBF_CHECK(e && e->have_state && out, SERBF_ERR_STATE, "norms: no state");
    SERBF_HIP(hipSetDevice(e->device));
    const int Nv = e->cfg.Nv, S = e->S;
    const size_t cols = (size_t)e->N * e->ntr;
    const double *ref = e->d_q0;
    DevTemp<double> tmp;
    if (q_ref) {
        SERBF_HIP(tmp.alloc(cols * S));
        SERBF_HIP(hipMemsetAsync(tmp.p, 0, cols * S * sizeof(double), e->stream));
        const int urc = upload_padded(e, tmp.p, q_ref, cols, Nv, S);
        if (urc) return urc;
        ref = tmp.p;
    }
    const int blocks = 1024;
    int rc = ensure_scratch(e, blocks * 10 * sizeof(double));
    if (rc) return rc;
    double *part = (double *)e->d_scratch;
    if (precision == SERBF_F32)
        hipLaunchKernelGGL(norms_kernel<float>, dim3(blocks), dim3(256), 0, e->stream, cols, Nv, S, ref, e->s32.qT,
                           e->d_wmass, part);
    else
        hipLaunchKernelGGL(norms_kernel<double>, dim3(blocks), dim3(256), 0, e->stream, cols, Nv, S, ref, e->s64.qT,
                           e->d_wmass, part);
    SERBF_HIP(hipGetLastError());
    std::vector<double> h(blocks * 10);
    SERBF_HIP(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    SERBF_HIP(hipStreamSynchronize(e->stream));
    double a[10] = {0, 0, 0, 0, 0, 0, 0, 0, -1e300, 1e300};
    for (int b = 0; b < blocks; b++) {
        for (int k = 0; k < 6; k++) a[k] += h[b * 10 + k];
        for (int k = 6; k < 9; k++) a[k] = std::fmax(a[k], h[b * 10 + k]);
        a[9] = std::fmin(a[9], h[b * 10 + 9]);
    }
    out[0] = a[0] / a[1];
    out[1] = std::sqrt(a[2] / a[3]);
    out[2] = a[6] / a[7];
    out[3] = a[4];
    out[4] = a[5];
    out[5] = a[9];
    out[6] = a[8];
    return SERBF_OK;
}

}  // extern "C"
