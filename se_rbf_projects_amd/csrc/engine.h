// Internal declarations of libserbf (MI355X RBF-FD tracer-advection engine).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/serbf.h"
#include "../../include/serbf_ref.h"

namespace serbf {

void set_error(const std::string &msg);

#define SERBF_HIP(call)                                                                   \
    do {                                                                                  \
        hipError_t err_ = (call);                                                         \
        if (err_ != hipSuccess) {                                                         \
            ::serbf::set_error(std::string(#call) + ": " + hipGetErrorString(err_));     \
            return SERBF_ERR_HIP;                                                         \
        }                                                                                 \
    } while (0)

#define SERBF_CHECK(cond, code, msg)                                                      \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            ::serbf::set_error(msg);                                                      \
            return code;                                                                  \
        }                                                                                 \
    } while (0)

// Device memory owned by one call: freed on every return path, early error returns included.
// hipFree waits for the device, so work still queued on the buffer finishes first.
template <typename T>
struct DevTemp {
    T *p = nullptr;
    DevTemp() = default;
    DevTemp(const DevTemp &) = delete;
    DevTemp &operator=(const DevTemp &) = delete;
    ~DevTemp()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t n) { return hipMalloc((void **)&p, n * sizeof(T)); }
};

// Largest vertical column handled by one wavefront (lane = level).
constexpr int kMaxLevels = 64;
// Largest vertical stencil.
constexpr int kMaxNv = 9;
// Largest horizontal stencil supported by the weight solver.
constexpr int kMaxStencil = 128;
// Per-node velocity basis record (see flows in rk4.hip).
constexpr int kBasis = 16;
// Scalars per stage time in the time table.
constexpr int kTimeScalars = 4;

// Device buffers of one precision of the time loop.
template <typename T>
struct StateBuffers {
    T *qT = nullptr;      // state at the start of the step   [N][ntr][S]
    T *qA = nullptr;      // stage states (ping-pong)
    T *qB = nullptr;
    T *F = nullptr;       // RK4 accumulator
    T *Wx = nullptr, *Wy = nullptr, *WL = nullptr;  // [N][n]
    T *basis = nullptr;   // [N][kBasis]
    T *lev = nullptr;     // [4][S]: r_inv, ga, gb, unused
    T *vW = nullptr;      // [S][kMaxNv]
    T *tscal = nullptr;   // [steps][4 stages][kTimeScalars]
    T *gamma = nullptr;   // per-point hyperviscosity coefficient [N][ntr][S] (closure)
    T *WL1 = nullptr, *WL2 = nullptr;  // Delta, Delta^2 weights [N][n] (closure features)
    _Float16 *s16 = nullptr;           // closure multiplier in FP16 [N][ntr][S] (FP32 tile kernel)
};

struct GraphCache {
    hipGraphExec_t exec = nullptr;
    int G = 0;
    int precision = -1;
    const void *key = nullptr;
    double dt = 0.0;
    int hv = -1;
};

}  // namespace serbf

struct serbf_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    int N = 0;
    int n = 0;
    // nodes
    double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
    double *d_lam = nullptr, *d_phi = nullptr;   // optional given longitude/latitude
    // stencils
    int *d_idx = nullptr;
    double *d_D = nullptr;
    bool have_stencils = false;
    // configuration
    serbf_config cfg{};
    bool configured = false;
    int S = 0;      // column stride = lanes per column
    int ntr = 1;    // tracers
    int nv = 0;     // vertical stencil size
    int *d_vstart = nullptr;
    int h_vstart[serbf::kMaxLevels];
    double h_vW[serbf::kMaxLevels * serbf::kMaxNv];
    // weights (FP64 master copy), solve status
    double *d_Wx = nullptr, *d_Wy = nullptr, *d_WL = nullptr;
    // config mode: Delta and Delta^2 weights (closure features), filled by the weight solves
    double *d_WL1 = nullptr, *d_WL2 = nullptr;
    bool have_lap = false;
    // per-node stencil sizes / polynomial degrees (empty: uniform), the weight-solve schedule
    std::vector<int> h_nn, h_deg;
    int *d_nn = nullptr, *d_deg = nullptr, *d_worder = nullptr;
    int wvariant = 4;            // schedule variant of the weight solves (autotuner; 4: register rows)
    double wstats_ms = 0.0, wstats_systems = 0.0;
    int wstats_variant = -1;
    int *d_status = nullptr;
    bool have_weights = false;
    // time loop buffers
    serbf::StateBuffers<double> s64;
    serbf::StateBuffers<float> s32;
    bool have_state = false;
    double *d_q0 = nullptr;   // FP64 copy of the state given to serbf_set_state
    double *d_wmass = nullptr;  // per-level rho(h) 4 pi r^2 dh / N
    bool f32_ready = false;   // FP32 weights/basis converted
    int tcap = 0;             // time table capacity in steps
    // [0] step counter read by the stage kernels (0 at the start of each run), [1] zero,
    // [2] the closure cadence origin: steps run since the state was set, modulo closure_every
    int *d_base = nullptr;
    long long steps_done = 0; // steps run since the state was last set
    long long closure_start = 0; // steps_done when the closure was attached
    // closure coefficient mode: 0 none/constant, 1 per-point gamma
    int hv_mode = 0;
    // hyperviscosity once per step (1) or in every stage (0)
    int hv_split = 0;
    // stage kernel: 7 vector tiles for FP32 (matrix-core tiles otherwise), 5 matrix-core tiles, 2 column form
    int kver = 7;
    // neighbour loads issued together per wavefront in the stage kernel (4, 8, 16)
    int unroll = 8;
    serbf::GraphCache graph;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // LDS tile plan of the stage kernel (16-node tiles): union ids, counts, local indices
    int *d_tuni = nullptr, *d_tcount = nullptr;
    unsigned short *d_lidx = nullptr;
    int tile_ustride = 0, tile_umax = 0, ntiles = 0;
    bool have_plan = false;
    // reusable scratch
    void *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    // learned closure attached to the time loop (not owned)
    serbf_closure *closure = nullptr;
    int closure_every = 8;
    double closure_smax = 8.0;
    double closure_kappa = 0.0;  // per-node stability bound of the multiplier (0: none)
    float *d_sbound = nullptr;   // [N] kappa G_ref / G_p
    float *d_hloc = nullptr;   // nearest-neighbour chord per node (closure features)
    float *d_nmax = nullptr;   // [N][4] per node and tracer max |q| (closure tile evaluation)
    int cus = 0;               // compute units of the device (0 until queried)
};

namespace serbf {
int ensure_scratch(serbf_engine *e, size_t bytes);
int alloc_stencils(serbf_engine *e, int n);
int alloc_weights(serbf_engine *e);
int convert_to_f32(serbf_engine *e, size_t n, const double *src, float *dst);
// kNN (knn.hip)
int knn_run(serbf_engine *e, int n);
// weights (weights.hip)
int weights_run(serbf_engine *e);
int weights_autotune(serbf_engine *e, int sample, double *report, int *nvariants);
int vertical_table(const serbf_config &cfg, int *start, double *W);
// time loop (rk4.hip)
int rk4_prepare(serbf_engine *e);
// LDS-tiled stage kernel (tile.hip)
int tile_plan(serbf_engine *e);
// Copies between the host layout [cols][Nv] and the device layout [cols][S] (levels padded to
// S with zeros), each as one contiguous transfer through a host staging buffer (a strided 2-D
// copy to or from pageable memory moves one row at a time).  Both return after the transfer.
int upload_padded(serbf_engine *e, double *dst, const double *src, size_t cols, int Nv, int S);
int download_padded(serbf_engine *e, double *dst, const double *src, size_t cols, int Nv, int S);

// matrix-core stage kernel (mtile.hip)
template <typename T>
struct StageArgs;
bool mtile_supported(const serbf_engine *e);
template <typename T>
int launch_mtile(serbf_engine *e, int stage, int hv, StageArgs<T> a);
void rk4_free(serbf_engine *e);
int rk4_rhs(serbf_engine *e, double t, double *F_host);
int rk4_run(serbf_engine *e, int precision, int nsteps, double dt, double t0, bool timed,
            double *ms, double *stage_ms);
int rk4_eval_array_velocity(serbf_engine *e, const double *Up_host, double alpha,
                            const double *q_host, double *F_host);
// the kTimeScalars flow time factors at time t (rk4.hip)
void time_scalars_public(const serbf_config &c, double t, double *out);
// closure (closure.hip): coefficients from the step-start state sb.qT into sb.gamma
template <typename T>
int closure_enqueue(serbf_engine *e, StateBuffers<T> &sb, double dt, int step_off, bool compact);
}  // namespace serbf
