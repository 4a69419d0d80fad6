// Host-side node sets and initial conditions of libserbf.
//
// Node sets: the reference reads its icosahedral sets from nodesets/ICOS_<N>.nc
// (nodesets.c:18-90); the two largest (163842, 655362) are not shipped, so they are
// generated here by recursive 4-way subdivision of the icosahedron with edge
// midpoints projected to the sphere.  Level L has N = 10*4^L + 2 nodes, the first
// 12 are the icosahedron vertices in the order of the reference files, and every
// level extends the previous one.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "../../include/serbf.h"
#include "../../include/serbf_ref.h"

namespace {

struct Vec3 {
    double x, y, z;
};

void icosahedron(std::vector<Vec3> &v, std::vector<std::array<int, 3>> &f)
{
    const double g = (1.0 + std::sqrt(5.0)) / 2.0;
    const double s = 1.0 / std::sqrt(1.0 + g * g);
    const double a = g * s, b = s;
    v = {{0, a, b},  {0, -a, b}, {0, a, -b}, {0, -a, -b}, {b, 0, a},  {-b, 0, a},
         {b, 0, -a}, {-b, 0, -a}, {a, b, 0}, {-a, b, 0}, {a, -b, 0}, {-a, -b, 0}};
    // faces: triples of mutually adjacent vertices (edge length 2b)
    auto adj = [&](int i, int j) {
        const double dx = v[i].x - v[j].x, dy = v[i].y - v[j].y, dz = v[i].z - v[j].z;
        return std::fabs(std::sqrt(dx * dx + dy * dy + dz * dz) - 2.0 * b) < 1e-9;
    };
    f.clear();
    for (int i = 0; i < 12; i++)
        for (int j = i + 1; j < 12; j++)
            for (int k = j + 1; k < 12; k++)
                if (adj(i, j) && adj(j, k) && adj(i, k)) f.push_back({i, j, k});
}

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double unif(uint64_t &s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// Hilbert index of (x, y) on a 2^order grid
uint64_t hilbert_d(uint32_t order, uint32_t x, uint32_t y)
{
    uint64_t d = 0;
    for (uint32_t s = 1u << (order - 1); s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1 : 0, ry = (y & s) ? 1 : 0;
        d += (uint64_t)s * s * ((3 * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - x;
                y = s - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

void lonlat1(double x, double y, double z, double *lam, double *phi)
{
    double l = std::atan2(y, x);
    if (l < 0.0) l += 2.0 * SERBF_PI;
    *lam = l;
    *phi = std::atan2(z, std::sqrt(x * x + y * y));
}

double chord_to(double x, double y, double z, double lc, double pc)
{
    const double cx = std::cos(pc) * std::cos(lc), cy = std::cos(pc) * std::sin(lc), cz = std::sin(pc);
    const double dx = x - cx, dy = y - cy, dz = z - cz;
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}
double angle_of_chord(double c) { return 2.0 * std::asin(std::fmin(1.0, 0.5 * c)); }

}  // namespace

extern "C" {

int serbf_icos_count(int level)
{
    if (level < 0 || level > 12) return -1;
    return 10 * (1 << (2 * level)) + 2;
}

int serbf_icos_nodes(int level, double *x, double *y, double *z)
{
    if (level < 0 || level > 12 || !x || !y || !z) return SERBF_ERR_ARG;
    std::vector<Vec3> v;
    std::vector<std::array<int, 3>> f;
    icosahedron(v, f);
    v.reserve(serbf_icos_count(level));
    for (int L = 0; L < level; L++) {
        std::unordered_map<uint64_t, int> mid;
        mid.reserve(f.size() * 2);
        std::vector<std::array<int, 3>> nf;
        nf.reserve(f.size() * 4);
        auto midpoint = [&](int a, int b) {
            const uint64_t key = ((uint64_t)std::min(a, b) << 32) | (uint64_t)std::max(a, b);
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            const Vec3 &p = v[a], &q = v[b];
            const double mx = p.x + q.x, my = p.y + q.y, mz = p.z + q.z;
            const double r = std::sqrt(mx * mx + my * my + mz * mz);
            v.push_back({mx / r, my / r, mz / r});
            const int id = (int)v.size() - 1;
            mid.emplace(key, id);
            return id;
        };
        for (const auto &t : f) {
            const int ab = midpoint(t[0], t[1]), bc = midpoint(t[1], t[2]), ca = midpoint(t[2], t[0]);
            nf.push_back({t[0], ab, ca});
            nf.push_back({t[1], bc, ab});
            nf.push_back({t[2], ca, bc});
            nf.push_back({ab, bc, ca});
        }
        f.swap(nf);
    }
    for (size_t i = 0; i < v.size(); i++) {
        x[i] = v[i].x;
        y[i] = v[i].y;
        z[i] = v[i].z;
    }
    return SERBF_OK;
}

int serbf_jitter_nodes(int N, double *x, double *y, double *z, double frac, unsigned long long seed)
{
    if (N <= 0 || !x || !y || !z) return SERBF_ERR_ARG;
    const double h = std::sqrt(8.0 * SERBF_PI / (std::sqrt(3.0) * N));
    uint64_t s = seed;
    for (int i = 0; i < N; i++) {
        const double px = x[i], py = y[i], pz = z[i];
        // orthonormal tangent basis: e1 = normalize(a x p) for the axis a least aligned with p
        double ax = 0, ay = 0, az = 0;
        if (std::fabs(px) <= std::fabs(py) && std::fabs(px) <= std::fabs(pz)) ax = 1;
        else if (std::fabs(py) <= std::fabs(pz)) ay = 1;
        else az = 1;
        double e1x = ay * pz - az * py, e1y = az * px - ax * pz, e1z = ax * py - ay * px;
        const double r1 = std::sqrt(e1x * e1x + e1y * e1y + e1z * e1z);
        e1x /= r1; e1y /= r1; e1z /= r1;
        const double e2x = py * e1z - pz * e1y, e2y = pz * e1x - px * e1z, e2z = px * e1y - py * e1x;
        const double d = frac * h * unif(s), th = 2.0 * SERBF_PI * unif(s);
        const double c = std::cos(th), sn = std::sin(th);
        const double qx = px + d * (c * e1x + sn * e2x), qy = py + d * (c * e1y + sn * e2y),
                     qz = pz + d * (c * e1z + sn * e2z);
        const double r = std::sqrt(qx * qx + qy * qy + qz * qz);
        x[i] = qx / r;
        y[i] = qy / r;
        z[i] = qz / r;
    }
    return SERBF_OK;
}

int serbf_spatial_order(int N, const double *x, const double *y, const double *z, int *perm)
{
    if (N <= 0 || !x || !y || !z || !perm) return SERBF_ERR_ARG;
    const uint32_t order = 16, side = 1u << order;
    std::vector<uint64_t> key(N);
    for (int i = 0; i < N; i++) {
        const double c[3] = {x[i], y[i], z[i]};
        int ax = 0;
        for (int k = 1; k < 3; k++)
            if (std::fabs(c[k]) > std::fabs(c[ax])) ax = k;
        const int face = ax + (c[ax] < 0 ? 3 : 0);
        const double m = std::fabs(c[ax]);
        const double u = c[(ax + 1) % 3] / m, w = c[(ax + 2) % 3] / m;
        auto q = [&](double t) {
            long v = (long)std::floor((t + 1.0) * 0.5 * side);
            return (uint32_t)std::min<long>(std::max<long>(v, 0), side - 1);
        };
        key[i] = ((uint64_t)face << 40) | hilbert_d(order, q(u), q(w));
    }
    std::vector<int> p(N);
    std::iota(p.begin(), p.end(), 0);
    std::stable_sort(p.begin(), p.end(), [&](int a, int b) { return key[a] < key[b]; });
    std::memcpy(perm, p.data(), N * sizeof(int));
    return SERBF_OK;
}

void serbf_lonlat(int N, const double *x, const double *y, const double *z, double *lambda, double *phi)
{
    for (int i = 0; i < N; i++) lonlat1(x[i], y[i], z[i], &lambda[i], &phi[i]);
}

// Initial tracer fields q[N*Nv] on cell-centred levels h_v = dh (v + 1/2).
int serbf_initial_condition(int kind, const serbf_config *cfg, int N, const double *x, const double *y,
                            const double *z, unsigned long long seed, double *q)
{
    if (!cfg || N <= 0 || !x || !y || !z || !q) return SERBF_ERR_ARG;
    const int Nv = cfg->Nv;
    const double dh = cfg->htop / Nv;
    double lc[2], pc[2];
    uint64_t s = seed;
    for (int k = 0; k < 2; k++) {
        lc[k] = 2.0 * SERBF_PI * unif(s);
        pc[k] = std::asin(2.0 * unif(s) - 1.0) * (2.0 / 3.0);
    }
    const double lcs[2] = {5.0 * SERBF_PI / 6.0, 7.0 * SERBF_PI / 6.0};
    for (int p = 0; p < N; p++) {
        double lam, phi;
        lonlat1(x[p], y[p], z[p], &lam, &phi);
        for (int v = 0; v < Nv; v++) {
            const double h = dh * (v + 0.5);
            double val = 1.0;
            if (kind == SERBF_IC_TC2) {
                // test_cases.c:171-180
                const double h1 = 2000.0, h2 = 5000.0, h0 = (h1 + h2) / 2;
                val = h < h1 || h > h2 ? 0.0 : (1.0 + std::cos((2.0 * SERBF_PI) * (h - h0) / (h2 - h1))) / 2.0;
            } else if (kind == SERBF_IC_BELL) {
                const double a = angle_of_chord(chord_to(x[p], y[p], z[p], 1.5 * SERBF_PI, 0.0));
                const double rb = 1.0 / 3.0;
                const double bell = a < rb ? 0.5 * (1.0 + std::cos(SERBF_PI * a / rb)) : 0.0;
                const double g = (h - 5000.0) / 1000.0;
                val = bell * std::exp(-0.5 * g * g);
            } else if (kind == SERBF_IC_HILLS) {
                const double g = (h - 0.5 * cfg->htop) / (0.2 * cfg->htop);
                double sum = 0.0;
                for (int k = 0; k < 2; k++) {
                    const double c = chord_to(x[p], y[p], z[p], lc[k], pc[k]) / 0.2;
                    sum += std::exp(-c * c);
                }
                val = sum * std::exp(-g * g);
            } else if (kind == SERBF_IC_SLOTTED) {
                const double r = 0.5;
                val = 0.1;
                for (int k = 0; k < 2; k++) {
                    const double a = angle_of_chord(chord_to(x[p], y[p], z[p], lcs[k], 0.0));
                    if (a <= r && (std::fabs(lam - lcs[k]) >= r / 6.0 || phi < -5.0 * r / 12.0)) val = 1.0;
                }
            } else if (kind == SERBF_IC_BELLS2) {
                const double r = 0.5;
                val = 0.1;
                for (int k = 0; k < 2; k++) {
                    const double a = angle_of_chord(chord_to(x[p], y[p], z[p], lcs[k], 0.0));
                    if (a < r) val += 0.9 * 0.5 * (1.0 + std::cos(SERBF_PI * a / r));
                }
            }
            q[(size_t)p * Nv + v] = val;
        }
    }
    return SERBF_OK;
}

}  // extern "C"
