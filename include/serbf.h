/*
 * serbf.h -- device-resident engine API of the MI355X RBF-FD tracer-advection
 * engine.  Plain C ABI: plain pointers and sizes, int status codes, no torch
 * or HIP types.  The reference-compatible drop-in symbols (get_stencils,
 * get_rbffd_DMs, get_solution, ...) are in serbf_ref.h and are built on top of
 * this API.
 *
 * Data conventions (host side):
 *   nodes      x, y, z double[N] on the unit sphere
 *   stencils   idx int32[N*n], D double[N*n], row-major, sorted by (d^2, index)
 *   weights    Dx, Dy, L double[N*n]: d/dx', d/dy' in the node's rotated frame
 *              (unit sphere) and the Delta^k operator (unit sphere)
 *   state      q double[N*ntracers*Nv], node-major ([node][tracer][level])
 */
#ifndef SERBF_H
#define SERBF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SERBF_OK 0
#define SERBF_ERR_ARG 1          /* invalid argument or size */
#define SERBF_ERR_HIP 2          /* HIP runtime error (message in serbf_last_error) */
#define SERBF_ERR_SINGULAR 3     /* at least one weight system was singular */
#define SERBF_ERR_UNSUPPORTED 4  /* configuration outside what the kernels support */
#define SERBF_ERR_STATE 5        /* call order: a prerequisite has not been set */
#define SERBF_ERR_NOMEM 6

/* method: how the weights and the vertical operator are built */
#define SERBF_MODE_REFERENCE 0   /* reference solver: PHS r^k + degree-5 tangent-plane
                                    monomials (unscaled), Gaussian+constant Delta^k system,
                                    7-point FD1 vertical with zero ghosts, H = I at the
                                    south pole (rotated vertical velocity sign kept) */
#define SERBF_MODE_CONFIG 1      /* PHS r^k + degree-d tangent-plane monomials on radius-
                                    scaled stencils, Gaussian or PHS Delta^k system, 1D RBF-FD vertical
                                    stencils (one-sided at the boundaries), radial
                                    velocity taken as w everywhere */

/* flows */
#define SERBF_FLOW_TC2 0         /* reference test case 2, Hadley-like overturning */
#define SERBF_FLOW_SOLID_BODY 1  /* tilted solid-body rotation, w = 0 */
#define SERBF_FLOW_DEFORM 2      /* solid-body rotation + reversing swirl + reversing w */

/* precision of the time loop */
#define SERBF_F64 0
#define SERBF_F32 1

typedef struct serbf_config {
    /* vertical geometry: cell-centred levels h_v = dh (v + 1/2), dh = htop / Nv */
    int Nv;
    double htop;            /* m */
    double R;               /* sphere radius, m */

    /* horizontal weights */
    int mode;               /* SERBF_MODE_* */
    int n;                  /* stencil size */
    int phs_order;          /* odd PHS order of the gradient system */
    int poly_deg;           /* tangent-plane polynomial degree (reference mode: 5) */
    int hv_order;           /* k of Delta^k (0: no hyperviscosity weights) */
    int hv_phs_order;       /* config mode: odd PHS order of the Delta^k system (>= 2k+1, with the
                               gradient system's monomials), or 0: Gaussian + constant Delta^k
                               system with shape parameter hv_eps (the reference's form) */
    double hv_eps;          /* Gaussian shape parameter epsilon of the unit sphere */

    /* vertical operator (config mode) */
    int nv;                 /* vertical stencil size (odd, <= 9) */
    int v_phs_order;        /* odd PHS order of the 1D system */
    int v_poly_deg;         /* 1D polynomial degree */

    /* right-hand side: dq/dt = -U . grad q + hv_coef * (Delta^k_unit q) */
    double hv_coef;         /* 1/s; 0 disables the term */

    /* flow */
    int flow;               /* SERBF_FLOW_* */
    double u0;              /* solid-body equatorial speed, m/s */
    double alpha;           /* solid-body tilt, rad */
    double T;               /* period of the reversing parts, s */
    double k_def;           /* swirl amplitude, m/s */
    double w0;              /* vertical velocity amplitude, m/s */

    /* tracers advected together (sharing stencils and weights); 0 means 1 */
    int ntracers;
} serbf_config;

typedef struct serbf_engine serbf_engine;

/* last error message of the calling thread */
const char *serbf_last_error(void);

/* ---- node sets (host) ---- */
/* Icosahedral node set from recursive 4-way subdivision: N = 10*4^level + 2.
 * Each level's set is a prefix of the next; the first 12 nodes are the icosahedron
 * vertices.  Arrays must hold N doubles. */
int serbf_icos_count(int level);
int serbf_icos_nodes(int level, double *x, double *y, double *z);
/* Seeded tangential jitter: each node moves by U(0, frac*h) in a uniformly random
 * tangent direction and is renormalised; h = sqrt(8 pi / (sqrt(3) N)). */
int serbf_jitter_nodes(int N, double *x, double *y, double *z, double frac, unsigned long long seed);
/* Node-set file (the reference's nodesets/<name>.nc files, HDF5 as netCDF-4 writes it), read without
 * netCDF: group "nodeset", float64 datasets x, y, z, lambda, phi.  Replaces the file reading of
 * get_nodeset_from_file (nodesets.c:18-90).  With x, y, z, lambda and phi all NULL, sets *N to
 * the node count only; otherwise *N is the capacity of the arrays on entry and the node count
 * on return (x, y, z required; lambda, phi may be NULL).  Contiguous datasets only;
 * anything else is refused with SERBF_ERR_ARG and a message.  Host only. */
int serbf_read_nodeset(const char *path, int *N, double *x, double *y, double *z, double *lambda, double *phi);
/* Permutation perm[new] = old that orders nodes along a space-filling curve
 * (Hilbert order of the cube-face coordinates) for gather locality. */
int serbf_spatial_order(int N, const double *x, const double *y, const double *z, int *perm);
/* lambda in [0, 2pi), phi = atan2(z, sqrt(x^2 + y^2)) */
void serbf_lonlat(int N, const double *x, const double *y, const double *z, double *lambda, double *phi);

/* initial tracer fields on cell-centred levels (q[N*Nv], node-major) */
#define SERBF_IC_TC2 0       /* reference test case 2: cos^2 layer between 2 and 5 km */
#define SERBF_IC_BELL 1      /* cosine bell of radius R/3 at (3pi/2, 0) x Gaussian in height
                                (centre 5 km, sigma 1 km) */
#define SERBF_IC_HILLS 2     /* two Gaussian hills of width 0.2R, centres from the seed,
                                x Gaussian in height (centre htop/2, width 0.2 htop) */
#define SERBF_IC_SLOTTED 3   /* two slotted cylinders of radius R/2 at (5pi/6, 0), (7pi/6, 0) */
#define SERBF_IC_BELLS2 4    /* two cosine bells of radius R/2 on a 0.1 background */
#define SERBF_IC_CONST 5     /* constant 1 (mass check) */
int serbf_initial_condition(int kind, const serbf_config *cfg, int N, const double *x, const double *y,
                            const double *z, unsigned long long seed, double *q);

/* ---- engine ---- */
/* *out is NULL on failure; nothing is left allocated. */
int serbf_create(serbf_engine **out, int device);
void serbf_destroy(serbf_engine *e);

int serbf_set_nodes(serbf_engine *e, int N, const double *x, const double *y, const double *z);
/* Optional longitude/latitude of the nodes (as read from a node file); without
 * them the flows use lambda = atan2(y, x) in [0, 2pi), phi = atan2(z, |(x, y)|). */
int serbf_set_lonlat(serbf_engine *e, const double *lambda, const double *phi);
/* GPU kNN (bucketed, exact (d^2, index) order). */
int serbf_knn(serbf_engine *e, int n);
int serbf_get_stencils(serbf_engine *e, int *idx, double *D);
int serbf_set_stencils(serbf_engine *e, int n, const int *idx);
/* The same kNN search (identical idx and D to serbf_knn + serbf_get_stencils) into caller-held
 * device memory, allocating nothing: no counterpart in the reference; it is for callers that
 * cannot allocate device memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  x, y, z (double[N]) and the outputs idx (int[N*n]) and D (double[N*n],
 * row-major, nearest first) are device memory on `device`; the search runs on that device's null
 * stream and returns when done.  Its working arrays are one fixed device array of 61 MiB per GPU,
 * held until the process ends, which bounds N to 1..1048576 (SERBF_ERR_ARG beyond).  Calls that
 * overlap on one GPU share that array and race. */
int serbf_knn_noalloc(int device, int N, const double *x, const double *y, const double *z, int n, int *idx,
                      double *D);

/* Drops weights and state.  On failure the engine is left unconfigured (configure again before
 * serbf_weights or serbf_set_state). */
int serbf_configure(serbf_engine *e, const serbf_config *cfg);
/* Batched FP64 weight solves for every node (one persistent launch). */
int serbf_weights(serbf_engine *e);
/* Per-node stencil sizes and polynomial degrees (config mode).  The stencil rows keep their
 * stride n (the largest size, from serbf_knn / serbf_set_stencils); node i uses its first
 * nn[i] neighbours (2 <= nn[i] <= n) and tangent-plane monomials of degree deg[i] (0..6, below
 * 2 hv_order), with nn[i] + (deg+1)(deg+2)/2 <= 112; its other row entries get zero weights, so
 * the time loop runs unchanged on the padded rows.  The Gaussian Delta^k shape parameter of a
 * node grows as sqrt(n / nn[i]), keeping eps times the stencil radius the same for every size.
 * All the uneven systems are solved in one persistent launch by serbf_weights.  nn = deg = NULL returns to uniform stencils.  Call after
 * serbf_configure. */
int serbf_set_stencil_sizes(serbf_engine *e, const int *nn, const int *deg);
/* Autotuner of the batched weight solves: times each schedule variant that fits on `sample`
 * evenly spaced nodes and keeps the fastest for serbf_weights.  Variants 0-3: the LDS kernel
 * (one wavefront per system, rows in LDS) with the refinement residual re-evaluated (0, 1) or
 * read from an LDS copy of the matrix (2, 3); 4-5 (4 the default): rows in registers (one or
 * two wavefronts per system with one row per thread up to 96 unknowns, four wavefronts with half
 * a row per thread up to 112; one persistent launch per size class and system kind); even
 * variants in node order, odd ones largest systems first.
 * report[3*v + {0,1,2}] = variant, ms (-1: does not fit), systems per second; *nvariants = 6
 * (report holds at least 48 doubles). */
int serbf_weights_autotune(serbf_engine *e, int sample, double *report, int *nvariants);
/* Pin the schedule variant of serbf_weights (0-5, as the autotuner numbers them). */
int serbf_weights_set_variant(serbf_engine *e, int variant);
/* The last serbf_weights launch: kernel time (ms), systems solved, variant used. */
int serbf_weights_stats(serbf_engine *e, double *ms, double *systems, int *variant);
/* status[i] = 0 ok, 1 singular gradient system, 2 singular Delta^k system */
int serbf_get_weights(serbf_engine *e, double *Dx, double *Dy, double *L, int *status);
int serbf_set_weights(serbf_engine *e, const double *Dx, const double *Dy, const double *L);
/* Vertical operator table: start[Nv], W[Nv*nv] (per-level stencil start and weights). */
int serbf_get_vertical(serbf_engine *e, int *start, double *W, int *nv);

int serbf_set_state(serbf_engine *e, const double *q);
int serbf_get_state(serbf_engine *e, int precision, double *q);
/* One right-hand-side evaluation F(q, t) of the FP64 state (for parity tests). */
int serbf_rhs(serbf_engine *e, double t, double *F);
/* nsteps classic RK4 steps from time t0 in the given precision.  The FP32 run
 * starts from the FP64 state converted once (serbf_set_state). */
int serbf_run(serbf_engine *e, int precision, int nsteps, double dt, double t0);
/* Same, timed with HIP events on the engine's stream; *ms = elapsed time of the
 * nsteps steps, *stage_ms = average duration of one stage kernel launch. */
int serbf_run_timed(serbf_engine *e, int precision, int nsteps, double dt, double t0,
                    double *ms, double *stage_ms);
/* Hyperviscosity placement: 0 (default) applies hv_coef * Delta^k q inside every RK4
 * stage (the right-hand side form); 1 applies it once per step as dt c Delta^k q(t)
 * from the step-start state (operator splitting), which removes the Delta^k stream
 * from stages 2-4. */
int serbf_set_hv_split(serbf_engine *e, int split);
/* Per-point hyperviscosity coefficients c[N*ntracers*Nv] (1/s) replacing hv_coef, e.g.
 * from the learned closure; NULL returns to the constant hv_coef. */
int serbf_set_gamma(serbf_engine *e, const double *gamma);
/* Stage kernel variant: 7 (default) as 5, with FP32 at 4 tracers on the vector tile kernel
 * (one wavefront per tracer, 16-byte state accesses, vertical stencils of at most 7 levels);
 * 5 matrix-core tiles with the state read directly as the matrix operand (FP32 and FP64, 1 or
 * 4 tracers of 64 levels, config-mode solid-body or deformational flow, hyperviscosity in
 * every stage; other cases use 2); 2 column-vector form (every shape). */
int serbf_set_kernel_version(serbf_engine *e, int version);
/* Neighbour columns each wavefront loads before consuming them (4, 8 or 16): the
 * memory-level parallelism / register trade-off of the stage kernel. */
int serbf_set_stage_unroll(serbf_engine *e, int unroll);
/* The tile plan of the matrix-core stage kernel (number of 16-node tiles, largest
 * neighbour union). */
int serbf_tile_info(serbf_engine *e, int *ntiles, int *umax);
/* Streams `bytes` from one device buffer to another reps times (16-byte vector
 * copy kernel); *ms_per_copy is the average launch time.  Reference for the
 * attainable HBM bandwidth and for calibrating the HBM byte counters. */
int serbf_stream_copy(serbf_engine *e, size_t bytes, int reps, double *ms_per_copy);
/* Delta and Delta^2 weights L1, L2 double[N*n] (config mode, poly_deg >= 2): solved together
 * with the gradient weights from the same PHS r^k + monomial saddle system (extra right-hand
 * sides); used by the closure features. */
int serbf_get_lap_weights(serbf_engine *e, double *L1, double *L2);

/* ---- learned hyperviscosity closure ----
 * A per-point multiplier s = min(softplus(MLP(x)), s_max, b_p) of the hyperviscosity
 * coefficient, c = hv_coef * s, with the MLP 10 -> 64 -> 64 -> 1 (ReLU hidden layers)
 * evaluated in BF16 on the matrix cores with FP32 accumulation.  s = 1 is the
 * constant-coefficient method, which is also the fallback wherever the features or s are not
 * finite.  s_max and the optional per-node stability bound b_p keep c inside the stability
 * region of the explicit RK4 step: b_p = kappa G_ref / G_p with G_p = sum_j |L_pj| the row sum
 * of the node's Delta^k weights (a Gershgorin radius of the hyperviscosity operator, which
 * scales as h^-8) and G_ref its 99.9th percentile over the nodes.
 *
 * Features of a point (node p, tracer, level v), with h the node's nearest-neighbour chord
 * (unit sphere) and asinh compression a(y) = asinh(8 y):
 *   x0 a(h |grad q|)   x1 a(h^2 Delta q)   x2 a(h^4 Delta^2 q)   x3 a(h^8 Delta^k q)
 *   x4 h / hbar (hbar = sqrt(8 pi / (sqrt(3) N)))   x5 |u| dt / (R h)   x6 |u| / u_max
 *   x7 (v + 1/2) / Nv   x8 q   x9 a(dh |dq/dz|)
 * Network input: (x_i - shift_i) * scale_i.
 * Parameters (float[SERBF_CLOSURE_PARAMS]): W1[64][10], b1[64], W2[64][64], b2[64], w3[64], b3. */
#define SERBF_CLOSURE_FEATURES 10
#define SERBF_CLOSURE_HIDDEN 64
#define SERBF_CLOSURE_PARAMS 4929
typedef struct serbf_closure serbf_closure;
/* *out is NULL on failure; nothing is left allocated. */
int serbf_closure_create(serbf_closure **out, int device);
void serbf_closure_destroy(serbf_closure *c);
/* params float[SERBF_CLOSURE_PARAMS]; norm float[32] = shift[16], scale[16] (first 10 used).
 * Resets the Adam moments. */
int serbf_closure_set_params(serbf_closure *c, const float *params, const float *norm);
int serbf_closure_get_params(serbf_closure *c, float *params);
/* The clamp s_max of the multiplier in the training loss (default 16, > 1). */
int serbf_closure_set_smax(serbf_closure *c, double s_max);
/* Training set: n samples of features feat[n][16] (first 10 used, raw as above), model
 * column A[n] and target R[n].  The loss of a sample is ((s - 1) A - R)^2, so s = 1 scores
 * mean(R^2).  Samples are used in the given order (shuffle beforehand). */
int serbf_closure_set_data(serbf_closure *c, long n, const float *feat, const float *A, const float *R);
/* Per-sample clamps cap[n] of the training loss, s = min(softplus(o), s_max, cap), for the
 * samples of the last serbf_closure_set_data (same n); cap = NULL removes them. */
int serbf_closure_set_caps(serbf_closure *c, long n, const float *cap);
/* Mean loss over samples [first, first + count) and its gradient: grad float[SERBF_CLOSURE_PARAMS + 1]
 * (the gradient, then the loss).  count must be a multiple of 16. */
int serbf_closure_grad(serbf_closure *c, long first, long count, float *grad);
/* steps Adam steps (beta 0.9, 0.999, eps 1e-8) over consecutive batches of `batch` samples
 * (a multiple of 16) of the range [first, first + count), continuing from the batch after the
 * last step taken and wrapping over the range; losses[steps] gets each batch's mean loss. */
int serbf_closure_train(serbf_closure *c, long first, long count, int steps, int batch, double lr,
                        float *losses);
/* Network output o (before the softplus) for samples [first, first + count), count % 16 == 0. */
int serbf_closure_eval(serbf_closure *c, long first, long count, float *o);
/* Features of the FP64 state at time t for time step dt: feat[(p*ntr + tr)*Nv + v][16] and
 * lap[...] = Delta^k q (unit sphere).  Needs config-mode weights with hv_order >= 3. */
int serbf_closure_features(serbf_engine *e, double t, double dt, float *feat, double *lap);
/* Use the closure in the engine's time loop: the coefficients are re-evaluated from the
 * step-start state every `every` steps (1..4096) counted from the first step after this
 * call, across serbf_run calls (runs of a and b steps refresh as one run of a + b steps;
 * setting the state restarts the count, and its first step evaluates), with the
 * multiplier clamped to s_max (0 < s_max <= 65504, the FP16 limit of the FP32 tile path).
 * c = NULL returns to the constant coefficient.  The engine keeps a pointer to c: detach
 * before destroying it. */
int serbf_closure_attach(serbf_engine *e, serbf_closure *c, int every, double s_max);
/* The per-node stability bound of the attached closure's multiplier: kappa > 0 sets
 * b_p = kappa G_ref / G_p from the current Delta^k weights (see above; recomputed when new
 * weights are attached), kappa = 0 removes it.  bound (optional, float[N]) receives b_p. */
int serbf_closure_bound(serbf_engine *e, double kappa, float *bound);
/* Evaluate the attached closure once on the state of the given precision at time t (step dt)
 * and return the coefficients gamma double[N*ntr*Nv] (1/s). */
int serbf_closure_gamma(serbf_engine *e, int precision, double t, double dt, double *gamma);

/* Reset the FP32 state from the FP64 state. */
int serbf_sync_f32_from_f64(serbf_engine *e);
/* Relative norms of state(precision) against q_ref (host, N*Nv):
 * out = {L1, L2, Linf, mass_ref, mass, min, max}; mass uses the reference's
 * rho(h) * 4 pi r^2 dh / N cell volumes. */
int serbf_norms(serbf_engine *e, int precision, const double *q_ref, double *out);

#ifdef __cplusplus
}
#endif
#endif /* SERBF_H */
