/*
 * ref_harness.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A driver (this file is ours) around the reference solver's own driver-level
 * sources, compiled unmodified from the reference tree by `make -C oracle ref`:
 *   CODE/wrappers.c   get_rt_config, init_phys_constants, init_global_layers,
 *                     init_local_patch, get_model_ICs, verify_solution
 *   CODE/test_cases.c init_TC2, set_TC2_U_t / update_U, TC2_u_*, q_init_TC2, HS_isoT_rho
 *   CODE/layers.c     get_layers
 *   CODE/patches.c    init_part_utils, init_patch_nodeset, init_patch_halos,
 *                     allocate_part_SV_data, allocate_patch_SV_data
 *   CODE/debug.c      abort_solver
 * (CODE = Advection/C_MPI_OMP_ACC/code).  The files that need MKL, netCDF or METIS
 * (rbffd_DMs.c, nodesets.c, domains.c) cannot be built in this image and are not
 * linked; the linker drops the wrappers.c functions that call them
 * (--gc-sections), and this driver never calls those.
 *
 * In place of init_global_nodeset (netCDF read + kNN) the driver loads the node set
 * and its stencils from files the caller wrote.  In place of partition_global_domain
 * it sets the one-part partition: METIS with one part gives every node part 0 and
 * the contiguous reorder is then the identity (domains.c:48-112), leaving
 * part_ids = 0, part_sizes = {N}, part_start_ids = {0}.
 *
 *   ref_harness <in_dir> <out_dir>     (ADV_* environment as the reference reads it)
 *
 * in_dir:  nodes.bin (x, y, z, lambda, phi; 5 x N f64), idx.bin (N x n i32),
 *          D.bin (N x n f64), times.bin (f64, optional), samples.bin ((h, phi) f64
 *          pairs, optional), qt.bin (N x Nv f64 state for verify_solution, optional).
 * out_dir: raw little-endian arrays named after the reference fields, plus
 *          scalars.txt ("name value" lines, doubles in %.17g).
 * verify_solution prints its banner to stdout.
 */
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "config.h"
#include "layers.h"
#include "patches.h"
#include "test_cases.h"
#include "wrappers.h"

/* main.c:10-11 */
int mpi_size;
int mpi_rank;

/* wrappers.c:23-31, test_cases.c:20 */
extern rt_config_struct rt_config[1];
extern phys_constants_struct phys_constants[1];
extern domains_struct global_domains[1];
extern layers_struct global_layers[1];
extern patch_struct local_patch[1];
extern TC2_params_struct TC2_params[1];

static const char *g_in, *g_out;

static void *load(const char *name, size_t *bytes, int required)
{
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s", g_in, name);
    FILE *f = fopen(path, "rb");
    if (!f) {
        if (required) {
            fprintf(stderr, "ref_harness: cannot open %s\n", path);
            exit(2);
        }
        *bytes = 0;
        return NULL;
    }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    void *p = malloc(sz > 0 ? (size_t)sz : 1);
    if (sz > 0 && fread(p, 1, (size_t)sz, f) != (size_t)sz) {
        fprintf(stderr, "ref_harness: short read %s\n", path);
        exit(2);
    }
    fclose(f);
    *bytes = (size_t)sz;
    return p;
}

static void dump(const char *name, const void *p, size_t bytes)
{
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s", g_out, name);
    FILE *f = fopen(path, "wb");
    if (!f || (bytes && fwrite(p, 1, bytes, f) != bytes)) {
        fprintf(stderr, "ref_harness: cannot write %s\n", path);
        exit(2);
    }
    fclose(f);
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: ref_harness <in_dir> <out_dir>\n");
        return 2;
    }
    g_in = argv[1];
    g_out = argv[2];
    MPI_Init(&argc, &argv);
    MPI_Comm_size(MPI_COMM_WORLD, &mpi_size);
    MPI_Comm_rank(MPI_COMM_WORLD, &mpi_rank);
    if (mpi_size != 1) {
        fprintf(stderr, "ref_harness: one rank only\n");
        MPI_Abort(MPI_COMM_WORLD, 2);
    }

    /* main.c:22-25 */
    get_rt_config();
    init_phys_constants();

    /* the global node set and its stencils (what init_global_nodeset produces) */
    size_t b_nodes, b_idx, b_D;
    double *nodes = load("nodes.bin", &b_nodes, 1);
    int N = (int)(b_nodes / (5 * sizeof(double)));
    int n = rt_config->stencil_size;
    int *idx = load("idx.bin", &b_idx, 1);
    double *D = load("D.bin", &b_D, 1);
    if (b_idx != (size_t)N * n * sizeof(int) || b_D != (size_t)N * n * sizeof(double)) {
        fprintf(stderr, "ref_harness: stencil files do not match N=%d n=%d\n", N, n);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    nodeset_struct *gn = global_domains->global_nodeset;
    gn->Nnodes = N;
    gn->n = n;
    gn->x = nodes;
    gn->y = nodes + N;
    gn->z = nodes + 2 * (size_t)N;
    gn->lambda = nodes + 3 * (size_t)N;
    gn->phi = nodes + 4 * (size_t)N;
    gn->idx = idx;
    gn->D = D;
    global_domains->Nnodes = N;

    /* main.c:35 */
    init_global_layers();

    /* the one-part partition (see the header) */
    global_domains->Nparts = 1;
    global_domains->part_ids = calloc((size_t)N, sizeof(int));
    global_domains->part_sizes = malloc(sizeof(int));
    global_domains->part_start_ids = malloc(sizeof(int));
    global_domains->part_sizes[0] = N;
    global_domains->part_start_ids[0] = 0;

    /* main.c:46, 60 */
    init_local_patch();
    get_model_ICs();

    /* ---- dump ---- */
    patch_struct *p = local_patch;
    layers_struct *ly = p->layers;
    const int Np = p->Nnodes, Npart = p->part_Nnodes, pNv = ly->padded_Nv, pNvt = ly->padded_Nvt;
    char sname[1024];
    snprintf(sname, sizeof(sname), "%s/scalars.txt", g_out);
    FILE *s = fopen(sname, "w");
    if (!s) {
        fprintf(stderr, "ref_harness: cannot write %s\n", sname);
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    fprintf(s, "rt.nodeset_from_file %d\nrt.nodeset_file %s\nrt.num_layers %d\nrt.model_height %.17g\n"
               "rt.stencil_size %d\nrt.TC %d\nrt.timestep_length %.17g\nrt.dt %.17g\nrt.nsteps %d\nrt.use_metis %d\n",
            rt_config->nodeset_from_file, rt_config->nodeset_file, rt_config->num_layers, rt_config->model_height,
            rt_config->stencil_size, rt_config->TC, rt_config->timestep_length, rt_config->dt, rt_config->nsteps,
            rt_config->use_metis);
    fprintf(s, "pc.R %.17g\npc.g %.17g\npc.p0 %.17g\npc.cp %.17g\npc.Rd %.17g\npc.kappa %.17g\npc.T0 %.17g\n"
               "pc.rho0 %.17g\npc.htop %.17g\n",
            phys_constants->R, phys_constants->g, phys_constants->p0, phys_constants->cp, phys_constants->Rd,
            phys_constants->kappa, phys_constants->T0, phys_constants->rho0, phys_constants->htop);
    fprintf(s, "tc2.tau %.17g\ntc2.K %.17g\ntc2.u0 %.17g\ntc2.w0 %.17g\ntc2.h1 %.17g\ntc2.h2 %.17g\n", TC2_params->tau,
            TC2_params->K, TC2_params->u0, TC2_params->w0, TC2_params->h1, TC2_params->h2);
    layers_struct *gl = global_layers;
    fprintf(s, "gl.Nv %d\ngl.Ng %d\ngl.padded_Nv %d\ngl.Nvt %d\ngl.padded_Nvt %d\ngl.dh %.17g\ngl.htop %.17g\n",
            gl->Nv, gl->Ng, gl->padded_Nv, gl->Nvt, gl->padded_Nvt, gl->dh, gl->htop);
    fprintf(s, "p.Nnodes %d\np.part_Nnodes %d\np.nodeset.Nnodes %d\np.nodeset.n %d\n", Np, Npart,
            p->nodeset->Nnodes, p->nodeset->n);
    fprintf(s, "p.halos.Nnbrs %d\np.halos.halo_size_sum %d\np.halos.nbr_halo_size_sum %d\n", p->halos->Nnbrs,
            p->halos->halo_size_sum, p->halos->nbr_halo_size_sum);
    fprintf(s, "p.SV_q_t.Ndim %d\np.SV_U.Ndim %d\np.SV_F.Ndim %d\np.SV_3D_temp.Ndim %d\n", p->SV_q_t->Ndim,
            p->SV_U->Ndim, p->SV_F->Ndim, p->SV_3D_temp->Ndim);
    fclose(s);

    dump("gl.h", gl->h, sizeof(double) * gl->padded_Nv);
    dump("gl.r", gl->r, sizeof(double) * gl->padded_Nv);
    dump("gl.r_inv", gl->r_inv, sizeof(double) * gl->padded_Nv);
    dump("p.layers.h", ly->h, sizeof(double) * pNv);
    dump("p.layers.r", ly->r, sizeof(double) * pNv);
    dump("p.layers.r_inv", ly->r_inv, sizeof(double) * pNv);
    dump("p.pid2gid", p->pid2gid, sizeof(int) * Np);
    dump("p.gid2pid", p->gid2pid, sizeof(int) * N);
    dump("p.part_pids", p->part_pids, sizeof(int) * Npart);
    dump("p.nodeset.x", p->nodeset->x, sizeof(double) * Np);
    dump("p.nodeset.y", p->nodeset->y, sizeof(double) * Np);
    dump("p.nodeset.z", p->nodeset->z, sizeof(double) * Np);
    dump("p.nodeset.lambda", p->nodeset->lambda, sizeof(double) * Np);
    dump("p.nodeset.phi", p->nodeset->phi, sizeof(double) * Np);
    dump("p.nodeset.idx", p->nodeset->idx, sizeof(int) * (size_t)Np * n);
    dump("p.nodeset.D", p->nodeset->D, sizeof(double) * (size_t)Np * n);
    /* patch arrays with their CACHE_ALIGN front pad (patches.c:335) */
    const size_t patch_len = (size_t)Np * pNvt + CACHE_ALIGN;
    dump("p.SV_q_init", p->SV_q_init->data - CACHE_ALIGN, sizeof(double) * patch_len);
    dump("p.SV_q_t", p->SV_q_t->data - CACHE_ALIGN, sizeof(double) * patch_len);
    dump("p.SV_q_k", p->SV_q_k->data - CACHE_ALIGN, sizeof(double) * patch_len);
    dump("p.SV_Usph", p->SV_Usph->data, sizeof(double) * (size_t)Npart * pNv * 3);
    dump("p.SV_U", p->SV_U->data, sizeof(double) * (size_t)Npart * pNv * 3);
    dump("p.SV_F", p->SV_F->data, sizeof(double) * (size_t)Npart * pNv);

    /* update_U at the requested times (test_cases.c:22-28, 106-148) */
    size_t b_t;
    double *times = load("times.bin", &b_t, 0);
    for (size_t k = 0; k < b_t / sizeof(double); k++) {
        update_U(p, times[k]);
        char name[64];
        snprintf(name, sizeof(name), "U.%zu", k);
        dump(name, p->SV_U->data, sizeof(double) * (size_t)Npart * pNv * 3);
    }

    /* the TC2 point functions (test_cases.c:151-228) at (h, phi) samples */
    size_t b_s;
    double *smp = load("samples.bin", &b_s, 0);
    const size_t ns = b_s / (2 * sizeof(double));
    if (ns) {
        double *o = malloc(sizeof(double) * 5 * ns);
        for (size_t k = 0; k < ns; k++) {
            const double h = smp[2 * k], ph = smp[2 * k + 1];
            o[5 * k + 0] = TC2_u_r(h, ph);
            o[5 * k + 1] = TC2_u_lambda(ph);
            o[5 * k + 2] = TC2_u_phi(h, ph);
            o[5 * k + 3] = q_init_TC2(h);
            o[5 * k + 4] = HS_isoT_rho(h);
        }
        dump("samples.out", o, sizeof(double) * 5 * ns);
        free(o);
    }

    /* verify_solution on a given state (wrappers.c:108-176) */
    size_t b_q;
    double *qt = load("qt.bin", &b_q, 0);
    if (qt) {
        const int Nv = ly->Nv;
        if (b_q != (size_t)Npart * Nv * sizeof(double)) {
            fprintf(stderr, "ref_harness: qt.bin does not match N x Nv\n");
            MPI_Abort(MPI_COMM_WORLD, 2);
        }
        for (int i = 0; i < Npart; i++)
            memcpy(p->SV_q_t->data + (size_t)p->part_pids[i] * pNvt, qt + (size_t)i * Nv, sizeof(double) * Nv);
        verify_solution();
    }
    fflush(stdout);
    MPI_Finalize();
    return 0;
}
