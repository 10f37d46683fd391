// Seeded synthetic inputs (host only).
//
// The reference seeds rand() with time(NULL) (src/main.cpp:24) and places
// particles by an O(N^2) random sequential addition on a 1000-point lattice
// per axis (src/cluster.cpp:53-113), so its initial state cannot be
// reproduced.  These generators take an explicit seed and use a counter-
// based SplitMix64 stream (Steele, Lea, Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014); the RSA uses a cell grid
// so that 10^5 spheres are placed in well under a second.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/fipi.h"

namespace {

inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// uniform double in [0,1) for (seed, counter)
inline double u01(uint64_t seed, uint64_t ctr)
{
    const uint64_t h = mix64(mix64(seed) ^ (ctr * 0xD1B54A32D192ED03ull));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

extern "C" {

int fipi_init_phi(int64_t g, double mean, double amp, uint64_t seed, double *phi)
{
    if (g < 0 || (g > 0 && !phi)) return FIPI_ERR_ARG;
    for (int64_t i = 0; i < g; i++) phi[i] = mean + amp * (2.0 * u01(seed, (uint64_t)i) - 1.0);
    return FIPI_OK;
}

int fipi_init_uniform(int64_t n, double lo, double hi, uint64_t seed, double *out)
{
    if (n < 0 || (n > 0 && !out)) return FIPI_ERR_ARG;
    for (int64_t i = 0; i < n; i++) out[i] = lo + (hi - lo) * u01(seed, (uint64_t)i);
    return FIPI_OK;
}

int64_t fipi_init_rsa(int64_t n, double r, double gap, double lx, double ly, double lz, uint64_t seed,
                      int64_t max_attempts, double *pos)
{
    // gap < 0 allows overlapping (soft-contact) packings down to 2r + gap > 0
    if (n < 0 || (n > 0 && !pos) || !(r > 0) || !(2.0 * r + gap > 0) || !(lx > 0 && ly > 0 && lz > 0) ||
        max_attempts < 1)
        return FIPI_ERR_ARG;
    const double dmin = 2.0 * r + gap;
    const double L[3] = {lx, ly, lz};
    int nc[3];
    double cs[3];
    for (int a = 0; a < 3; a++) {
        nc[a] = (int)std::floor(L[a] / dmin);
        if (nc[a] < 1) nc[a] = 1;
        cs[a] = L[a] / nc[a];
    }
    const int64_t ncell = (int64_t)nc[0] * nc[1] * nc[2];
    std::vector<int64_t> head(ncell, -1), next(n > 0 ? n : 1, -1);
    uint64_t ctr = 0;
    int64_t placed = 0;
    for (int64_t i = 0; i < n; i++) {
        bool ok = false;
        double p[3];
        for (int64_t att = 0; att < max_attempts && !ok; att++) {
            for (int a = 0; a < 3; a++) {
                p[a] = L[a] * u01(seed, ctr++);
                if (p[a] >= L[a]) p[a] = 0.0;
            }
            int ci[3];
            for (int a = 0; a < 3; a++) {
                ci[a] = (int)std::floor(p[a] / cs[a]);
                if (ci[a] >= nc[a]) ci[a] = nc[a] - 1;
            }
            ok = true;
            const int rx = nc[0] >= 3 ? 1 : nc[0] - 1, ry = nc[1] >= 3 ? 1 : nc[1] - 1, rz = nc[2] >= 3 ? 1 : nc[2] - 1;
            for (int dz = -rz; dz <= rz && ok; dz++)
                for (int dy = -ry; dy <= ry && ok; dy++)
                    for (int dx = -rx; dx <= rx && ok; dx++) {
                        const int cx = (ci[0] + dx + nc[0]) % nc[0], cy = (ci[1] + dy + nc[1]) % nc[1],
                                  cz = (ci[2] + dz + nc[2]) % nc[2];
                        for (int64_t j = head[cx + (int64_t)nc[0] * (cy + (int64_t)nc[1] * cz)]; j >= 0; j = next[j]) {
                            double s = 0.0;
                            for (int a = 0; a < 3; a++) {
                                double d = p[a] - pos[3 * j + a];
                                d -= L[a] * std::nearbyint(d / L[a]);
                                s += d * d;
                            }
                            if (s < dmin * dmin) { ok = false; break; }
                        }
                    }
            if (ok) {
                for (int a = 0; a < 3; a++) pos[3 * i + a] = p[a];
                const int64_t c = ci[0] + (int64_t)nc[0] * (ci[1] + (int64_t)nc[1] * ci[2]);
                next[i] = head[c];
                head[c] = i;
            }
        }
        if (!ok) break;
        placed++;
    }
    return placed;
}

}  // extern "C"
