/*
 * oracle.h -- CPU oracle of the RBF-FD tracer-advection hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C (C11 + OpenMP) restatement of the
 * reference solver's algorithms, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product library (libserbf.so)
 * never links, loads or calls it.
 *
 * Two families of functions:
 *   orc_ref_*  restate the reference C solver function by function on its own
 *              structs (serbf_ref.h): get_stencils, init_rbffd_HH_rot_Ms,
 *              get_rbffd_DMs, eval_partSVDotGradSV, the RK4 axpys, update_Up,
 *              set_TC2_U_t, get_solution, verify_solution.
 *   orc_*      the generic method on flat arrays (serbf_config): reference mode
 *              and config mode weights, vertical operators, flows, right-hand
 *              side and RK4, norms.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#ifndef SERBF_ORACLE_H
#define SERBF_ORACLE_H

#include "../include/serbf.h"
#include "../include/serbf_ref.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- a1: kNN (nodesets.c:113-268) ---- */
void orc_knn_brute(int N, const double *x, const double *y, const double *z, int n,
                   int *idx, double *D);

/* ---- a2: Householder (rbffd_DMs.c:55-96); fix_pole != 0 maps the south pole
 *      with diag(1,-1,-1) instead of the reference's identity ---- */
void orc_householder(double x, double y, double z, int fix_pole, double H[9]);

/* ---- dense LU with partial pivoting, LAPACK dgesv semantics on row-major
 *      storage: returns 0 or k+1 for the first zero pivot ---- */
int orc_dgesv(int m, int nrhs, double *A, int *ipiv, double *B);
int orc_lu_factor(int m, double *A, int *ipiv);
void orc_lu_solve(int m, int nrhs, const double *LU, const int *ipiv, double *B);

/* ---- a3: weights ---- */
/* Reference mode, one node (rbffd_DMs.c:156-261): B3 gets size x 3 rotated-frame
 * weights (PHS r^k_phs + degree-5 monomials), BL gets n+1 Gaussian Delta^k_L weights. */
int orc_ref_node_weights(int n, const double *X /* 3 x n */, const double H[9], int k_phs,
                         int k_L, double eps, double *hDxp, double *hDyp, double *Lw,
                         int *info2);
/* Generic: all nodes, either mode. Dx, Dy, L are N*n; status[i] as serbf_get_weights. */
int orc_weights(const serbf_config *cfg, int N, const double *x, const double *y,
                const double *z, const int *idx, double *Dx, double *Dy, double *L,
                int *status);
/* Same, config mode also returning the Delta (L1) and Delta^2 (L2) weights solved with the
 * gradient system (PHS r^k + the same monomials); L1/L2 may be NULL. */
int orc_weights_lap(const serbf_config *cfg, int N, const double *x, const double *y,
                    const double *z, const int *idx, double *Dx, double *Dy, double *L,
                    double *L1, double *L2, int *status);
/* Config mode with per-node stencil size nn[i] <= n and polynomial degree deg[i]; rows keep
 * the stride n, entries past nn[i] are zero. */
int orc_weights_var(const serbf_config *cfg, int N, const double *x, const double *y,
                    const double *z, const int *idx, const int *nn, const int *deg, double *Dx,
                    double *Dy, double *L, int *status);
/* Vertical operator: start[Nv], W[Nv*nv]; returns nv. */
int orc_vertical(const serbf_config *cfg, int *start, double *W);

/* ---- flows: rotated-frame velocity Up (3 values) at node (x,y,z), height h ---- */
void orc_velocity(const serbf_config *cfg, double x, double y, double z, double h, double t,
                  double Up[3]);
/* Cartesian velocity U at the node */
void orc_velocity_cart(const serbf_config *cfg, double x, double y, double z, double h,
                       double t, double U[3]);

/* ---- initial conditions, q[N*Nv] ---- */
#define ORC_IC_TC2 0          /* cos^2 layer between 2 and 5 km */
#define ORC_IC_BELL 1         /* cosine bell R/3 x Gaussian in height (sigma 1 km, 5 km) */
#define ORC_IC_HILLS 2        /* two Gaussian hills (width 0.2R) x Gaussian in height */
#define ORC_IC_SLOTTED 3      /* slotted cylinder */
#define ORC_IC_BELLS2 4       /* two cosine bells */
#define ORC_IC_CONST 5        /* constant 1 */
void orc_initial(int kind, const serbf_config *cfg, int N, const double *x, const double *y,
                 const double *z, unsigned long long seed, double *q);

/* ---- right-hand side and RK4 on q[N*Nv] (node-major) ---- */
void orc_rhs(const serbf_config *cfg, int N, const double *x, const double *y, const double *z,
             const int *idx, const double *Dx, const double *Dy, const double *L,
             const int *vstart, const double *vW, const double *gamma /* N*Nv or NULL */,
             const double *q, double t, double *F);
void orc_rk4(const serbf_config *cfg, int N, const double *x, const double *y, const double *z,
             const int *idx, const double *Dx, const double *Dy, const double *L,
             const int *vstart, const double *vW, int nsteps, double dt, double t0,
             double *q);

/* OpenMP team size of the parallel loops */
void orc_set_threads(int n);
int orc_max_threads(void);

/* {L1, L2, Linf, mass_ref, mass, min, max}; Linf denominator max|q| (wrappers.c:152) */
void orc_norms(const serbf_config *cfg, int N, const double *q_ref, const double *q,
               double out[7]);

/* ---- struct-level restatements of the reference functions ---- */
void orc_ref_get_stencils(nodeset_struct *ns, int n);
void orc_ref_get_layers(layers_struct *layers, double htop, int Nv, double R);
void orc_ref_init_HH(patch_struct *p);
void orc_ref_get_rbffd_DMs(patch_struct *p, int k_phs, int k_L, double eps, double R);
void orc_ref_eval_partSVDotGradSV(patch_struct *p, double alpha, part_SV_struct *A,
                                  patch_SV_struct *B, part_SV_struct *C);
void orc_ref_eval_sumpatchSVpartSV(patch_struct *p, patch_SV_struct *A, double alpha,
                                   part_SV_struct *B, patch_SV_struct *C);
void orc_ref_eval_sumpartSVpartSV(patch_struct *p, part_SV_struct *A, double alpha,
                                  part_SV_struct *B, part_SV_struct *C);
void orc_ref_update_Up(patch_struct *p);
void orc_ref_set_TC2_U_t(patch_struct *p, double t, const phys_constants_struct *pc);
void orc_ref_init_TC2(patch_struct *p, rt_config_struct *rt, const phys_constants_struct *pc);
void orc_ref_get_solution(patch_struct *p, const rt_config_struct *rt,
                          const phys_constants_struct *pc);
void orc_ref_verify(patch_struct *p, int global_N, const phys_constants_struct *pc,
                    double out[7]);

/* ---- learned closure (closure.c): network, loss gradient, Adam, features ---- */
#define ORC_CLOSURE_PARAMS 4929
float orc_bf16(float x);
double orc_mlp_forward(const float *params, const float *norm, const float *feat, int emulate,
                       double *h1, double *h2);
double orc_mlp_grad(const float *params, const float *norm, long ns, const float *feat,
                    const float *A, const float *R, const float *cap, int emulate, double s_max,
                    double *grad);
void orc_adam(int np, float *params, double *m, double *v, const double *g, int step, double lr,
              double b1, double b2, double eps);
void orc_closure_features(const serbf_config *cfg, int N, const double *x, const double *y,
                          const double *z, const int *idx, const double *Dx, const double *Dy,
                          const double *L, const double *L1, const double *L2, const int *vstart,
                          const double *vW, int nv, int ntr, const double *q, double t, double dt,
                          float *feat, double *lap);

#ifdef __cplusplus
}
#endif
#endif
