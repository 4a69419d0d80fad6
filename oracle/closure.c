/*
 * closure.c -- oracle of the learned hyperviscosity closure.  TEST INFRASTRUCTURE ONLY.
 *
 * The closure has no counterpart in the reference (the reference applies a constant
 * hyperviscosity coefficient, PY/Solver.ipynb:428-430); this file restates the network
 * (10 -> 64 -> 64 -> 1, ReLU hidden layers, softplus output), its loss with the per-sample
 * clamps, the analytic gradient, Adam, and the per-point features, in double precision.  With emulate != 0 the
 * forward and backward passes round the operands of every matrix product to bfloat16
 * (round to nearest even) at the same points as the matrix-core kernels, and nowhere else.
 *
 * Parameter layout: W1[64][10], b1[64], W2[64][64], b2[64], w3[64], b3 (4929 floats).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NF 10
#define NH 64
#define OB1 (NF * NH)
#define OW2 (OB1 + NH)
#define OB2 (OW2 + NH * NH)
#define OW3 (OB2 + NH)
#define OB3 (OW3 + NH)

float orc_bf16(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    memcpy(&r, &u, 4);
    return r;
}

static double rb(double x, int emulate) { return emulate ? (double)orc_bf16((float)x) : x; }

static double softplus(double o) { return o > 20.0 ? o : log1p(exp(o)); }
static double relu(double z) { return z > 0.0 ? z : 0.0; }
static double sigmoid(double o) { return 1.0 / (1.0 + exp(-o)); }

/* network input of a sample: (x_f - shift_f) * scale_f, as float (the kernels normalise in FP32) */
static void inputs(const float *norm, const float *feat, int emulate, double x[NF])
{
    for (int f = 0; f < NF; f++) {
        const float v = (feat[f] - norm[f]) * norm[16 + f];
        x[f] = rb((double)v, emulate);
    }
}

double orc_mlp_forward(const float *P, const float *norm, const float *feat, int emulate, double *h1,
                       double *h2)
{
    double x[NF], a[NH], b[NH];
    double *H1 = h1 ? h1 : a, *H2 = h2 ? h2 : b;
    inputs(norm, feat, emulate, x);
    for (int h = 0; h < NH; h++) {
        double z = 0.0;
        for (int f = 0; f < NF; f++)
            z += rb(P[h * NF + f], emulate) * x[f];
        H1[h] = relu((double)(float)z + P[OB1 + h]);
    }
    for (int n = 0; n < NH; n++) {
        double z = 0.0;
        for (int h = 0; h < NH; h++)
            z += rb(P[OW2 + n * NH + h], emulate) * rb(H1[h], emulate);
        H2[n] = relu((double)(float)z + P[OB2 + n]);
    }
    double o = P[OB3];
    for (int n = 0; n < NH; n++)
        o += P[OW3 + n] * H2[n];
    return o;
}

double orc_mlp_grad(const float *P, const float *norm, long ns, const float *feat, const float *A,
                    const float *R, const float *cap, int emulate, double s_max, double *grad)
{
    memset(grad, 0, sizeof(double) * ORC_CLOSURE_PARAMS);
    double loss = 0.0;
    double x[NF], h1[NH], h2[NH], dz2[NH], dz1[NH];
    for (long s = 0; s < ns; s++) {
        const float *fs = feat + s * 16;
        const double o = orc_mlp_forward(P, norm, fs, emulate, h1, h2);
        inputs(norm, fs, emulate, x);
        /* s = min(softplus(o), s_max, cap[s]), no gradient through the clamp (s_max <= 0 and
         * cap == NULL: no clamp); the clamp is compared in float, as the kernels do */
        double smx = s_max;
        if (cap && (smx <= 0.0 || (double)cap[s] < smx))
            smx = (double)cap[s];
        const double sp = softplus(o);
        const int clamped = smx > 0.0 && sp > smx;
        const double e = ((clamped ? smx : sp) - 1.0) * A[s] - R[s];
        loss += e * e;
        const double go = clamped ? 0.0 : 2.0 * e * A[s] * sigmoid(o) / (double)ns;
        grad[OB3] += go;
        for (int n = 0; n < NH; n++) {
            grad[OW3 + n] += go * h2[n];
            dz2[n] = h2[n] > 0.0 ? go * P[OW3 + n] : 0.0;
            grad[OB2 + n] += dz2[n];
        }
        for (int n = 0; n < NH; n++)
            for (int h = 0; h < NH; h++)
                grad[OW2 + n * NH + h] += rb(dz2[n], emulate) * rb(h1[h], emulate);
        for (int h = 0; h < NH; h++) {
            double d = 0.0;
            for (int n = 0; n < NH; n++)
                d += rb(P[OW2 + n * NH + h], emulate) * rb(dz2[n], emulate);
            dz1[h] = h1[h] > 0.0 ? d : 0.0;
            grad[OB1 + h] += dz1[h];
        }
        for (int h = 0; h < NH; h++)
            for (int f = 0; f < NF; f++)
                grad[h * NF + f] += rb(dz1[h], emulate) * x[f];
    }
    return loss / (double)ns;
}

void orc_adam(int np, float *P, double *m, double *v, const double *g, int step, double lr,
              double b1, double b2, double eps)
{
    const double c1 = 1.0 - pow(b1, step), c2 = 1.0 - pow(b2, step);
    for (int k = 0; k < np; k++) {
        m[k] = b1 * m[k] + (1.0 - b1) * g[k];
        v[k] = b2 * v[k] + (1.0 - b2) * g[k] * g[k];
        P[k] = (float)(P[k] - lr * (m[k] / c1) / (sqrt(v[k] / c2) + eps));
    }
}

/* Features of every point (node p, tracer tr, level v < Nv) of the state q[N][ntr][Nv] at time
 * t (step dt): feat[((p*ntr + tr)*Nv + v)*16 + k] and lap[...] = Delta^k q; see serbf.h for the
 * definitions.  Dx, Dy, L, L1, L2 are the config-mode weights [N][n]; vstart/vW the vertical
 * operator [Nv], [Nv][nv]. */
void orc_closure_features(const serbf_config *cfg, int N, const double *x, const double *y,
                          const double *z, const int *idx, const double *Dx, const double *Dy,
                          const double *L, const double *L1, const double *L2, const int *vstart,
                          const double *vW, int nv, int ntr, const double *q, double t, double dt,
                          float *feat, double *lap)
{
    const int n = cfg->n, Nv = cfg->Nv;
    const double K = 8.0;
    const double hbar = sqrt(8.0 * SERBF_PI / (sqrt(3.0) * N));
    double umax = cfg->flow == SERBF_FLOW_DEFORM ? cfg->u0 + cfg->k_def : cfg->u0;
    if (!(umax > 0.0))
        umax = 1.0;
    const double dh = cfg->htop / Nv;
#pragma omp parallel for schedule(static)
    for (int p = 0; p < N; p++) {
        const int nb1 = idx[(size_t)p * n + (n > 1 ? 1 : 0)];
        const double ex = x[nb1] - x[p], ey = y[nb1] - y[p], ez = z[nb1] - z[p];
        const double h = (double)(float)sqrt(ex * ex + ey * ey + ez * ez);
        const double h2 = h * h, h4 = h2 * h2, h8 = h4 * h4;
        for (int tr = 0; tr < ntr; tr++) {
            for (int v = 0; v < Nv; v++) {
                double gx = 0, gy = 0, l1 = 0, l2 = 0, l4 = 0;
                for (int k = 0; k < n; k++) {
                    const size_t o = (size_t)p * n + k;
                    const double qv = q[((size_t)idx[o] * ntr + tr) * Nv + v];
                    gx += Dx[o] * qv;
                    gy += Dy[o] * qv;
                    l1 += L1[o] * qv;
                    l2 += L2[o] * qv;
                    l4 += L[o] * qv;
                }
                const double *own = q + ((size_t)p * ntr + tr) * Nv;
                double dzv = 0.0;
                for (int j = 0; j < nv; j++) {
                    const int k = vstart[v] + j;
                    if (k >= 0 && k < Nv)
                        dzv += vW[v * nv + j] * own[k];
                }
                const double hv = dh * (v + 0.5);
                double U[3];
                orc_velocity(cfg, x[p], y[p], z[p], hv, t, U);
                const double sp = sqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
                float *f = feat + (((size_t)p * ntr + tr) * Nv + v) * 16;
                f[0] = (float)asinh(K * h * sqrt(gx * gx + gy * gy));
                f[1] = (float)asinh(K * h2 * l1);
                f[2] = (float)asinh(K * h4 * l2);
                f[3] = (float)asinh(K * h8 * l4);
                f[4] = (float)(h / hbar);
                f[5] = (float)(sp * dt / (cfg->R * h));
                f[6] = (float)(sp / umax);
                f[7] = (float)((v + 0.5) / Nv);
                f[8] = (float)own[v];
                f[9] = (float)asinh(K * dh * fabs(dzv));
                for (int k = NF; k < 16; k++)
                    f[k] = 0.0f;
                lap[((size_t)p * ntr + tr) * Nv + v] = l4;
            }
        }
    }
}
