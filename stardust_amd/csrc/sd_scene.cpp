// Seeded synthetic scene generators (host only).
//
// The reference builds scenes from unseeded rand() (sandbox/testDEM.cpp:53-58,90-98)
// and lattice entities (src/engine/entities/entity.cpp:9-59). These generators give
// reproducible inputs for every benchmark configuration instead:
//   SplitMix64 stream, uniform = (u >> 40) * 2^-24, normal = Box-Muller,
//   lognormal = exp(ln(median) + sigma * normal), clipped to [r_min, r_max],
//   random sequential addition (RSA) on a hash grid so no two spheres overlap,
//   largest spheres first (ids keep their radius, so id order stays unrelated
//   to position; big-first addition reaches far higher packing than random order).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "sd_internal.hpp"

namespace sd {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uniform() { return (float)(next() >> 40) * (1.0f / 16777216.0f); }
    double normal() {
        const double u1 = 1.0 - (double)(next() >> 11) * (1.0 / 9007199254740992.0);  // (0, 1]
        const double u2 = (double)(next() >> 11) * (1.0 / 9007199254740992.0);
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

}  // namespace sd

extern "C" sd_status sd_gen_spheres(uint64_t seed, int n, const float lo[3], const float hi[3], float r_median,
                                    float r_sigma, float r_min, float r_max, float* x, float* y, float* z,
                                    float* radius) {
    SD_REQUIRE(n >= 0 && lo && hi && x && y && z && radius, "sd_gen_spheres: bad argument");
    SD_REQUIRE(r_median > 0.0f && r_sigma >= 0.0f && r_min > 0.0f && r_max >= r_min,
               "sd_gen_spheres: bad radius distribution");
    for (int a = 0; a < 3; ++a)
        SD_REQUIRE(hi[a] - lo[a] > 2.0f * r_max, "sd_gen_spheres: domain smaller than a sphere");
    sd::SplitMix64 rng(seed);
    // radii first (one value per sphere, in id order)
    for (int i = 0; i < n; ++i) {
        float r = r_median;
        if (r_sigma > 0.0f) r = (float)exp(log((double)r_median) + (double)r_sigma * rng.normal());
        r = r < r_min ? r_min : (r > r_max ? r_max : r);
        radius[i] = r;
    }
    // RSA on a grid of cell 2 r_max: an overlap partner lies in the 27 cells.
    const double cell = 2.0 * (double)r_max;
    int nc[3];
    for (int a = 0; a < 3; ++a) {
        double c = ceil(((double)hi[a] - (double)lo[a]) / cell);
        SD_REQUIRE(c < 4096.0, "sd_gen_spheres: grid too large");
        nc[a] = (int)c < 1 ? 1 : (int)c;
    }
    const size_t ncells = (size_t)nc[0] * nc[1] * nc[2];
    std::vector<int> head(ncells, -1), next((size_t)n, -1);
    auto cell_of = [&](float p, int a) {
        int c = (int)floor(((double)p - (double)lo[a]) / cell);
        return c < 0 ? 0 : (c >= nc[a] ? nc[a] - 1 : c);
    };
    const int max_attempts = 20000;
    std::vector<int> order((size_t)n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return radius[a] > radius[b]; });
    for (int oi = 0; oi < n; ++oi) {
        const int i = order[(size_t)oi];
        const float r = radius[i];
        bool placed = false;
        for (int att = 0; att < max_attempts && !placed; ++att) {
            float p[3];
            for (int a = 0; a < 3; ++a) p[a] = (lo[a] + r) + rng.uniform() * ((hi[a] - r) - (lo[a] + r));
            const int cx = cell_of(p[0], 0), cy = cell_of(p[1], 1), cz = cell_of(p[2], 2);
            bool ok = true;
            for (int dz = -1; dz <= 1 && ok; ++dz) {
                const int zz = cz + dz;
                if (zz < 0 || zz >= nc[2]) continue;
                for (int dy = -1; dy <= 1 && ok; ++dy) {
                    const int yy = cy + dy;
                    if (yy < 0 || yy >= nc[1]) continue;
                    for (int dx = -1; dx <= 1 && ok; ++dx) {
                        const int xx = cx + dx;
                        if (xx < 0 || xx >= nc[0]) continue;
                        for (int j = head[((size_t)zz * nc[1] + yy) * nc[0] + xx]; j >= 0; j = next[j]) {
                            const double ddx = (double)p[0] - x[j], ddy = (double)p[1] - y[j],
                                         ddz = (double)p[2] - z[j];
                            const double rs = (double)r + (double)radius[j];
                            if (ddx * ddx + ddy * ddy + ddz * ddz < rs * rs) {
                                ok = false;
                                break;
                            }
                        }
                    }
                }
            }
            if (ok) {
                x[i] = p[0];
                y[i] = p[1];
                z[i] = p[2];
                const size_t c = ((size_t)cz * nc[1] + cy) * nc[0] + cx;
                next[i] = head[c];
                head[c] = i;
                placed = true;
            }
        }
        if (!placed)
            return sd::set_error(SD_ERR_CAPACITY, "sd_gen_spheres: could not place sphere %d of %d without overlap",
                                 i, n);
    }
    return SD_OK;
}

// Rigid clumps. Per body, in this order: sphere count U{c_min..c_max}; per
// sphere a centre N(0, sigma) per axis and a radius U[r_min, r_max]; an
// orientation (Shoemake); a lattice jitter U[-jitter, jitter] per axis; a
// velocity N(0, v_sigma) per axis. Sphere masses rho 4/3 pi r^3; the body sits
// at its jittered lattice site with the member centres re-expressed about
// their centre of mass; inertia about the centre of mass is the sum of the
// spheres' 2/5 m r^2 plus the parallel-axis terms (replaces the per-entity
// packing of EntityHandler::prepareArrays, entity_handler.cpp:61-94).
extern "C" sd_status sd_gen_clumps(uint64_t seed, int nb, int lattice, float spacing, float jitter, int c_min,
                                   int c_max, float sigma, float r_min, float r_max, float rho, float v_sigma,
                                   float* pos4, float* vel4, float* quat4, float* angvel4, float* inertia6,
                                   int* first, int* count, float* local4, int* ns) {
    SD_REQUIRE(nb >= 1 && lattice >= 1 && (long)lattice * lattice * lattice >= nb, "sd_gen_clumps: bad lattice");
    SD_REQUIRE(c_min >= 1 && c_max >= c_min && r_min > 0.0f && r_max >= r_min && rho > 0.0f,
               "sd_gen_clumps: bad clump parameters");
    SD_REQUIRE(pos4 && vel4 && quat4 && angvel4 && inertia6 && first && count && local4 && ns,
               "sd_gen_clumps: NULL array");
    sd::SplitMix64 rng(seed);
    int total = 0;
    std::vector<double> c, r, m;
    for (int b = 0; b < nb; ++b) {
        const int k = c_min + (int)(rng.next() % (uint64_t)(c_max - c_min + 1));
        c.assign(3 * (size_t)k, 0.0);
        r.assign(k, 0.0);
        m.assign(k, 0.0);
        double M = 0.0, com[3] = {0, 0, 0};
        for (int s = 0; s < k; ++s) {
            for (int a = 0; a < 3; ++a) c[3 * s + a] = (double)sigma * rng.normal();
            r[s] = (double)r_min + (double)rng.uniform() * ((double)r_max - (double)r_min);
            m[s] = (double)rho * 4.0 / 3.0 * 3.141592653589793 * r[s] * r[s] * r[s];
            M += m[s];
            for (int a = 0; a < 3; ++a) com[a] += m[s] * c[3 * s + a];
        }
        for (int a = 0; a < 3; ++a) com[a] /= M;
        double I[6] = {0, 0, 0, 0, 0, 0};  // xx yy zz xy xz yz
        first[b] = total;
        count[b] = k;
        for (int s = 0; s < k; ++s) {
            const double d[3] = {c[3 * s] - com[0], c[3 * s + 1] - com[1], c[3 * s + 2] - com[2]};
            const double own = 0.4 * m[s] * r[s] * r[s];
            const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            I[0] += own + m[s] * (dd - d[0] * d[0]);
            I[1] += own + m[s] * (dd - d[1] * d[1]);
            I[2] += own + m[s] * (dd - d[2] * d[2]);
            I[3] += -m[s] * d[0] * d[1];
            I[4] += -m[s] * d[0] * d[2];
            I[5] += -m[s] * d[1] * d[2];
            float* l = local4 + 4 * (size_t)(total + s);
            for (int a = 0; a < 3; ++a) l[a] = (float)d[a];
            l[3] = (float)r[s];
        }
        total += k;
        for (int a = 0; a < 6; ++a) inertia6[6 * (size_t)b + a] = (float)I[a];
        const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        const double s1 = sqrt(1.0 - u1), s2 = sqrt(u1);
        const double t1 = 6.283185307179586 * u2, t2 = 6.283185307179586 * u3;
        float* q = quat4 + 4 * (size_t)b;
        q[0] = (float)(s2 * cos(t2));
        q[1] = (float)(s1 * sin(t1));
        q[2] = (float)(s1 * cos(t1));
        q[3] = (float)(s2 * sin(t2));
        const int li = b % lattice, lj = (b / lattice) % lattice, lk = b / (lattice * lattice);
        const int idx[3] = {li, lj, lk};
        float* p = pos4 + 4 * (size_t)b;
        for (int a = 0; a < 3; ++a)
            p[a] = (float)(((double)idx[a] + 0.5) * (double)spacing +
                           (double)jitter * (2.0 * (double)rng.uniform() - 1.0));
        p[3] = (float)M;
        float* v = vel4 + 4 * (size_t)b;
        for (int a = 0; a < 3; ++a) v[a] = (float)((double)v_sigma * rng.normal());
        v[3] = 0.0f;
        float* w = angvel4 + 4 * (size_t)b;
        w[0] = w[1] = w[2] = w[3] = 0.0f;
    }
    *ns = total;
    return SD_OK;
}

// Convex hulls of points on ellipsoids. Every point of a strictly convex
// surface is a vertex of the hull of the sample, so the point set is the hull.
// Per hull, in this order: vertex count U{nv_min..nv_max}, semi-axes U[ax_min,
// ax_max] (x, y, z), the vertices (Gaussian 3-vectors normalised and scaled by
// the semi-axes), an orientation (Shoemake's uniform unit quaternion, stored
// (w, x, y, z)), a centre uniform in [lo + ax_max, hi - ax_max] and a velocity
// N(0, v_sigma) per component.
extern "C" sd_status sd_gen_hulls(uint64_t seed, int n, const float lo[3], const float hi[3], int nv_min, int nv_max,
                                  float ax_min, float ax_max, float v_sigma, float* pos4, float* quat4, float* vel4,
                                  int* vstart, int* nv, float* verts4, int verts_capacity, int* n_verts) {
    SD_REQUIRE(n >= 0 && lo && hi && pos4 && quat4 && vel4 && vstart && nv && verts4 && n_verts,
               "sd_gen_hulls: bad argument");
    SD_REQUIRE(nv_min >= 1 && nv_max >= nv_min && nv_max <= 1024, "sd_gen_hulls: vertex counts must be in [1, 1024]");
    SD_REQUIRE(ax_min > 0.0f && ax_max >= ax_min, "sd_gen_hulls: bad semi-axes");
    sd::SplitMix64 rng(seed);
    int total = 0;
    for (int i = 0; i < n; ++i) {
        const int k = nv_min + (int)(rng.next() % (uint64_t)(nv_max - nv_min + 1));
        if (total + k > verts_capacity)
            return sd::set_error(SD_ERR_CAPACITY, "sd_gen_hulls: vertex capacity %d exceeded", verts_capacity);
        float ax[3];
        for (int a = 0; a < 3; ++a) ax[a] = ax_min + rng.uniform() * (ax_max - ax_min);
        vstart[i] = total;
        nv[i] = k;
        for (int j = 0; j < k; ++j) {
            double g[3], l2 = 0.0;
            do {
                l2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    g[a] = rng.normal();
                    l2 += g[a] * g[a];
                }
            } while (l2 < 1e-12);
            const double il = 1.0 / sqrt(l2);
            float* v = verts4 + 4 * (size_t)(total + j);
            for (int a = 0; a < 3; ++a) v[a] = (float)(g[a] * il * (double)ax[a]);
            v[3] = 0.0f;
        }
        total += k;
        const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        const double s1 = sqrt(1.0 - u1), s2 = sqrt(u1);
        const double t1 = 6.283185307179586 * u2, t2 = 6.283185307179586 * u3;
        float* q = quat4 + 4 * (size_t)i;
        q[0] = (float)(s2 * cos(t2));  // w
        q[1] = (float)(s1 * sin(t1));  // x
        q[2] = (float)(s1 * cos(t1));  // y
        q[3] = (float)(s2 * sin(t2));  // z
        float* p = pos4 + 4 * (size_t)i;
        for (int a = 0; a < 3; ++a) p[a] = (lo[a] + ax_max) + rng.uniform() * ((hi[a] - ax_max) - (lo[a] + ax_max));
        p[3] = 0.0f;
        float* v = vel4 + 4 * (size_t)i;
        for (int a = 0; a < 3; ++a) v[a] = (float)(v_sigma * rng.normal());
        v[3] = 0.0f;
    }
    *n_verts = total;
    return SD_OK;
}
