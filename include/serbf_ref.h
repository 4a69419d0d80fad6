/*
 * serbf_ref.h -- reference-compatible C interface of the RBF-FD tracer-advection
 * engine (single process, single GPU, no MPI).
 *
 * The struct layouts below are binary-compatible with the reference solver's
 * headers (Advection/C_MPI_OMP_ACC/code/include/config.h), so a caller compiled
 * against those headers can pass its structs straight to this library.  The only
 * change is that the two MPI handle pointers of patch_SV_struct become void*,
 * which keeps every size and offset on LP64.  Every function declared here is a
 * drop-in for the reference function cited next to it; libserbf.so runs its body
 * on the GPU (hand-written gfx950 kernels) and writes results back into the
 * caller's host structs with the reference's ownership rules (the callee
 * allocates output arrays with malloc/calloc and stores the pointers).
 */
#ifndef SERBF_REF_H
#define SERBF_REF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Compile-time constants of the reference (config.h:12-31). */
#define SERBF_REF_POLY_ORDER 5          /* 2D tangent-plane monomials up to degree 5 -> 21 terms */
#define SERBF_REF_NPOLY(d) ((((d) + 1) * ((d) + 2)) / 2)
#define SERBF_REF_PHS_ORDER 5           /* PHS r^5 */
#define SERBF_REF_MAX_HV_ORDER 16
#define SERBF_REF_HV_ORDER 4            /* Delta^4 hyperviscosity */
#define SERBF_REF_GHOST 3
#define SERBF_REF_FD1_SIZE 7
#define SERBF_REF_CACHE_ALIGN 8         /* doubles of front pad / column alignment */
#define SERBF_ALIGNUP(n, m) (((((n) - 1) / (m)) + 1) * (m))

/* Physical constants (config.h:46-52). */
#define SERBF_EARTH_RADIUS 6.37122e6
#define SERBF_GRAVITY 9.80616e0
#define SERBF_P0 1.0e5
#define SERBF_CP 1004.5
#define SERBF_RD 287.0
#define SERBF_T0 300.0
#define SERBF_PI 3.141592653589793

/* config.h:58-70 */
typedef struct phys_constants_struct {
    double R, g, p0, cp, Rd, kappa, T0, rho0, htop;
} phys_constants_struct;

/* config.h:96-120 (MAX_PATH_SIZE = 200) */
typedef struct rt_config_struct {
    int nodeset_from_file;
    char nodeset_file[200];
    int num_layers;
    double model_height;
    int stencil_size;
    int TC;
    double timestep_length;
    double dt;
    int nsteps;
    int use_metis;
} rt_config_struct;

/* config.h:129-174.  idx/D are row-major Nnodes x n, ascending distance per row. */
typedef struct nodeset_struct {
    int Nnodes;
    int n;
    double *x, *y, *z;          /* unit-sphere Cartesian coordinates */
    double *lambda, *phi;       /* longitude in [0, 2pi), latitude */
    int *idx;
    double *D;
} nodeset_struct;

/* config.h:180-211 */
typedef struct layers_struct {
    int Nv, Ng, padded_Nv, Nvt, padded_Nvt;
    double dh, htop;
    double *h, *r, *r_inv;
    double *p, *rho;
} layers_struct;

/* config.h:217-255 */
typedef struct domains_struct {
    int Nnodes, Nparts;
    nodeset_struct global_nodeset[1];
    int *part_ids, *part_sizes, *part_start_ids;
} domains_struct;

/* config.h:261-289 */
typedef struct halos_struct {
    int Nnbrs, halo_size_sum, nbr_halo_size_sum;
    int *nbr_ranks, *halo_sizes, *halo_offsets, *nbr_halo_sizes, *nbr_halo_offsets;
    int *nbr_hid2pid, *nbr_hid2gid, *hid2pid, *hid2gid;
} halos_struct;

/* config.h:292-318 */
typedef struct rbffds_DMs_struct {
    int part_Nnodes;
    int n;
    double gamma;
    int *idx;
    double *D;
    double *H;                              /* part_Nnodes x 9, row-major 3x3 */
    double *hDx, *hDy, *hDz;                /* Cartesian horizontal weights */
    double *vDx, *vDy, *vDz;                /* Cartesian vertical weights, part_Nnodes x 7 */
    double *L;                              /* Delta^k weights / R */
    double *hDxp, *hDyp;                    /* rotated-frame d/dx', d/dy' weights */
    double vDzp[SERBF_REF_FD1_SIZE];        /* unit FD1 weights / dh */
} rbffd_DMs_struct;

/* config.h:321-333; the MPI_Request* / MPI_Status* members are void* here. */
typedef struct patch_SV_struct {
    int Ndim;
    double *data;                           /* column stride padded_Nvt, +CACHE_ALIGN front pad */
    double *halo_buff, *nbr_halo_buff;
    void *request;
    void *status;
} patch_SV_struct;

/* config.h:335-340 */
typedef struct part_SV_struct {
    int Ndim;
    double *data;                           /* (partid*Ndim + dim)*padded_Nv + level */
} part_SV_struct;

/* config.h:347-390 */
typedef struct patch_struct {
    int Nnodes;
    int part_Nnodes;
    nodeset_struct nodeset[1];
    layers_struct layers[1];
    rbffd_DMs_struct rbffd_DMs[1];
    int *pid2gid, *gid2pid;
    int *part_pids;
    halos_struct halos[1];
    patch_SV_struct *SV_q_init, *SV_q_t, *SV_q_k;
    part_SV_struct *SV_U, *SV_Up, *SV_Usph;
    part_SV_struct *SV_F, *SV_F_k;
    part_SV_struct *SV_3D_temp;
} patch_struct;

/* test_cases.h:6-14 */
typedef struct TC2_params {
    double tau, K, u0, w0, h1, h2, htop;
} TC2_params_struct;

/* ---- Drop-in entry points (same names and signatures as the reference) ---- */

/* nodesets.h / nodesets.c:18-90: reads a node-set file (x, y, z, lambda, phi of group
 * "nodeset"), mallocs the five arrays into *nodeset, sets Nnodes and returns it.  Reads the HDF5
 * file directly, without netCDF or MPI (one process).  A file it cannot read ends the process
 * with a message, as the library's other drop-in entry points do. */
int get_nodeset_from_file(char *nodeset_file, nodeset_struct *nodeset);

/* nodesets.h:8 / nodesets.c:113-153: n nearest neighbours of every node by chord
 * distance, sorted by (d^2, index); allocates nodeset->idx and nodeset->D.
 * Runs the bucketed GPU kNN. */
void get_stencils(nodeset_struct *nodeset, int n);

/* layers.h:6 / layers.c:15-57: cell-centred levels dh = htop/Nv. */
void get_layers(layers_struct *layers, double htop, int Nv);

/* rbffd_DMs.h:6 / rbffd_DMs.c:18-53: copy partition rows of idx/D. */
void get_part_rbffd_stencils(patch_struct *local_patch);

/* rbffd_DMs.h:8 / rbffd_DMs.c:55-96: per-node Householder reflector H (H x0 = e_z). */
void init_rbffd_HH_rot_Ms(patch_struct *local_patch);

/* rbffd_DMs.h:10 / rbffd_DMs.c:98-282: all per-node weight solves, batched on
 * the GPU in one launch (FP64 LU with partial pivoting and refinement). */
void get_rbffd_DMs(patch_struct *local_patch, int k_phs, int k_L, double ep);

/* rbffd_DMs.h:12-20 / rbffd_DMs.c:284-328: scalar kernels. */
double phsrbf(double r, int k);
double d1_phsrbf(double r, double ri, int k);
double garbf(double r, double epsilon);
double L_garbf(double r, double epsilon, int k);
double L2_norm(double x, double y, double z);

/* rbffd_DMs.h:22 / rbffd_DMs.c:331-443: C = alpha * A . grad(B) (stencil SpMV + RHS). */
void eval_partSVDotGradSV(patch_struct *local_patch, double alpha, part_SV_struct *SV_A,
                          patch_SV_struct *SV_B, part_SV_struct *SV_C);
/* rbffd_DMs.h:24-28 / rbffd_DMs.c:446-520: C = A + alpha*B with patch/partition strides. */
void eval_sumpatchSVpartSV(patch_struct *local_patch, patch_SV_struct *SV_A, double alpha,
                           part_SV_struct *SV_B, patch_SV_struct *SV_C);
void eval_sumpartSVpartSV(patch_struct *local_patch, part_SV_struct *SV_A, double alpha,
                          part_SV_struct *SV_B, part_SV_struct *SV_C);
void eval_sumpatchSVpatchSV(patch_struct *local_patch, patch_SV_struct *SV_A, double alpha,
                            patch_SV_struct *SV_B, part_SV_struct *SV_C);
/* rbffd_DMs.h:30 / rbffd_DMs.c:523-544: Up = H U per node. */
void update_Up(patch_struct *local_patch);

/* test_cases.h:18-32 / test_cases.c:22-229: test case 2 (Hadley-like flow). */
void update_U(patch_struct *local_patch, double t);
void init_TC2(patch_struct *local_patch);
double q_init_TC2(double h);
double HS_isoT_rho(double h);
void set_TC2_U_t(patch_struct *local_patch, double t);
double TC2_u_r(double h, double phi);
double TC2_u_lambda(double phi);
double TC2_u_phi(double h, double phi);

/* wrappers.h:6-24 / wrappers.c:39-215: driver level.  get_solution runs the whole
 * RK4 time loop device-resident (captured as a HIP graph) on the global
 * local_patch and downloads q_t at the end. */
void get_solution(void);
void verify_solution(void);
void init_patch_rbffd_DMs(void);
void init_phys_constants(void);
/* wrappers.h:10 / wrappers.c:346-395: rt_config from the ADV_* environment with the
 * defaults of config.h:83-89, and the same parameter table on stdout (no broadcast). */
void get_rt_config(void);
/* wrappers.h:24 / wrappers.c:274-298: global_domains->global_nodeset from
 * rt_config->nodeset_file, then its stencils (GPU kNN). */
void init_global_nodeset(void);
/* wrappers.h:18 / wrappers.c:260-268: global_layers from rt_config. */
void init_global_layers(void);
/* wrappers.h:26 / wrappers.c:304-322: one process, one part: every node in part 0 and
 * the node order kept, as METIS and the contiguous reorder (domains.c:48-112) give at
 * one part.  mpi_size other than 1 ends the process with a message. */
void partition_global_domain(void);
/* wrappers.h:20 / wrappers.c:222-254 with patches.c:24-346 at one part: local_patch's
 * node set, maps, empty halos, layers and zero-filled state arrays (patch arrays with
 * stride padded_Nvt after a CACHE_ALIGN front pad, partition arrays with padded_Nv). */
void init_local_patch(void);
/* wrappers.h:8 / wrappers.c:183-193: init_TC2(local_patch). */
void get_model_ICs(void);

/* Globals the reference defines in wrappers.c:23-31 and main.c:10-11. */
extern rt_config_struct rt_config[1];
extern phys_constants_struct phys_constants[1];
extern domains_struct global_domains[1];
extern layers_struct global_layers[1];
extern patch_struct local_patch[1];
extern TC2_params_struct TC2_params[1];
extern int mpi_rank;
extern int mpi_size;

/* Verification results of the last verify_solution() call:
 * {L1, L2, Linf, mass_real, mass_approx, min, max} (wrappers.c:163-171). */
void serbf_ref_last_verification(double out[7]);

#ifdef __cplusplus
}
#endif
#endif /* SERBF_REF_H */
