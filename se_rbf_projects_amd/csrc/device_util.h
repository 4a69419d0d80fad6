// Device helpers shared by the kernels of libserbf.
#pragma once
#include <hip/hip_runtime.h>

namespace serbf {

// Householder reflector H = I - beta v v^T with H x0 = e_z (rbffd_DMs.c:75-91 of the
// reference).  x^2 + y^2 == 0 gives H = I there; fix_pole replaces that case at the
// south pole by diag(1, -1, -1) so that H x0 = +e_z holds at every node.
__device__ inline void householder_dev(double x, double y, double z, bool fix_pole, double H[9])
{
    const double dot_lon = x * x + y * y;
    double v0 = x, v1 = y, v2 = 1.0, beta;
    if (dot_lon == 0.0) {
        beta = 0.0;
    } else {
        v2 = z > 0.0 ? -dot_lon / (z + 1.0) : z - 1.0;
        beta = 2.0 * (v2 * v2) / (dot_lon + (v2 * v2));
        v0 = v0 / v2;
        v1 = v1 / v2;
        v2 = 1.0;
    }
    const double v[3] = {v0, v1, v2};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H[3 * i + j] = (i == j ? 1.0 : 0.0) + (-beta * v[i]) * v[j];
    if (fix_pole && dot_lon == 0.0 && z < 0.0) {
        for (int i = 0; i < 9; i++) H[i] = 0.0;
        H[0] = 1.0;
        H[4] = -1.0;
        H[8] = -1.0;
    }
}

}  // namespace serbf
