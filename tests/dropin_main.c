/*
 * dropin_main.c -- the reference solver's driver sequence (CODE/main.c:22-67, with
 * CODE = Advection/C_MPI_OMP_ACC/code) as a plain C program linked against libserbf.so
 * alone: the one-process form of main.c without MPI_Init / MPI_Finalize.  Test
 * infrastructure for the drop-in boundary; a user's own main.c calls the same
 * functions in the same order.
 *
 *   ADV_NODESET_FILE=<file.nc> ADV_NUM_LAYERS=.. ADV_STENCIL_SIZE=.. ... dropin_main [q_out]
 *
 * With q_out it also writes the final q_t (part_Nnodes x Nv doubles, node-major).
 */
#include <stdio.h>
#include <string.h>
#include <time.h>

#include "serbf_ref.h"

/* main.c:10-11 */
int mpi_size = 1;
int mpi_rank = 0;

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv)
{
    double tstart;
    get_rt_config();
    init_phys_constants();

    tstart = now();
    init_global_nodeset();
    printf("\ninit_global_nodeset() walltime -> %.2f seconds\n", now() - tstart);
    init_global_layers();

    tstart = now();
    partition_global_domain();
    printf("\npartition_global_domain() walltime -> %.2f seconds\n", now() - tstart);

    tstart = now();
    init_local_patch();
    printf("\ninit_local_patch() walltime -> %.2f seconds\n", now() - tstart);

    tstart = now();
    init_patch_rbffd_DMs();
    printf("\ninit_patch_rbffd_DMs() walltime -> %.2f seconds\n", now() - tstart);

    tstart = now();
    get_model_ICs();
    printf("\nget_model_ICs() walltime -> %.2f seconds\n", now() - tstart);

    get_solution();
    verify_solution();

    if (argc > 1) {
        FILE *f = fopen(argv[1], "wb");
        if (!f) {
            fprintf(stderr, "dropin_main: cannot write %s\n", argv[1]);
            return 2;
        }
        const patch_struct *p = local_patch;
        const int Nv = p->layers->Nv, pNvt = p->layers->padded_Nvt;
        for (int i = 0; i < p->part_Nnodes; i++)
            fwrite(p->SV_q_t->data + (size_t)p->part_pids[i] * pNvt, sizeof(double), (size_t)Nv, f);
        fclose(f);
    }
    return 0;
}
