/*
 * Oracle kNN (test infrastructure only).
 *
 * Restates get_stencils / get_DD2_slice / get_D_idx_slice
 * (Advection/C_MPI_OMP_ACC/code/nodesets.c:113-268):
 *   d^2 = (xi-xj)^2 + (yi-yj)^2 + (zi-zj)^2 summed left to right (nodesets.c:178-180,
 *         pow(.,2) is an exact square), no FMA contraction (-ffp-contract=off);
 *   insertion into a list initialised to D = 5.0 / idx = -1 (nodesets.c:210-214)
 *   with a strict '>' (nodesets.c:233), scanning candidates j in increasing order,
 *   so equal distances keep the lower index first and the result is the n smallest
 *   (d^2, j) keys in ascending order;
 *   D = sqrt(d^2) (nodesets.c:258-267).
 * The insertion here walks from the tail of the list, which yields the same list
 * as the reference's walk from the head.
 */
#include <math.h>
#include <stdlib.h>
#include "oracle.h"

void orc_knn_brute(int N, const double *x, const double *y, const double *z, int n,
                   int *idx, double *D)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < N; i++) {
        double *bd = D + (size_t)i * n;
        int *bi = idx + (size_t)i * n;
        for (int k = 0; k < n; k++) {
            bd[k] = 5.0;
            bi[k] = -1;
        }
        const double xi = x[i], yi = y[i], zi = z[i];
        for (int j = 0; j < N; j++) {
            double dx = xi - x[j], dy = yi - y[j], dz = zi - z[j];
            double d = dx * dx;
            d = d + dy * dy;
            d = d + dz * dz;
            if (!(bd[n - 1] > d))
                continue;
            int k = n - 1;
            while (k > 0 && bd[k - 1] > d) {
                bd[k] = bd[k - 1];
                bi[k] = bi[k - 1];
                k--;
            }
            bd[k] = d;
            bi[k] = j;
        }
        for (int k = 0; k < n; k++)
            bd[k] = sqrt(bd[k]);
    }
}
