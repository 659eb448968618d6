// Host-side generator of the synthetic periodic structures used by the
// batched fitting workload: random sequential addition with a minimum
// distance over every periodic image.  No device code; it lives in the
// library so that large batches are generated in seconds, identically on
// every machine (splitmix64 stream per structure).
#include <cmath>
#include <cstdint>
#include <vector>

#include "agpu_internal.hpp"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

using namespace agpu;

extern "C" int agpu_random_positions(int n, const double *cell, double dmin, uint64_t seed, int max_tries,
                                     double *pos)
{
    if (n < 0 || !cell || (n > 0 && !pos) || !(dmin >= 0.0) || max_tries < 1)
        return set_error(AGPU_ERR_ARG, "random_positions: bad arguments");
    double B[9];
    if (compute_inverse(cell, B) != AGPU_OK) return set_error(AGPU_ERR_ARG, "random_positions: singular cell");
    // After d -= round(d) every fractional difference is in [-1/2, 1/2]; an
    // image n_k can only be within dmin if |d_k + n_k| <= dmin |b_k| (b_k the
    // reciprocal vectors, b_k[x] = B[3x + k]), so |n_k| <= 1/2 + dmin |b_k|.
    int ext[3];
    for (int k = 0; k < 3; ++k) {
        double bk = std::sqrt(B[k] * B[k] + B[3 + k] * B[3 + k] + B[6 + k] * B[6 + k]);
        ext[k] = (int)std::floor(0.5 + dmin * bk) + 1;
    }
    std::vector<double> sh;   // Cartesian translations
    for (int a = -ext[0]; a <= ext[0]; ++a)
        for (int b = -ext[1]; b <= ext[1]; ++b)
            for (int c = -ext[2]; c <= ext[2]; ++c)
                for (int x = 0; x < 3; ++x) sh.push_back(a * cell[x] + b * cell[3 + x] + c * cell[6 + x]);
    const int ns = (int)sh.size() / 3;
    std::vector<double> fr(3 * (size_t)n);
    SplitMix rng{seed};
    double scale = 1.0;
    int m = 0, stall = 0;
    while (m < n) {
        const double d2min = dmin * dmin * scale;
        bool self_ok = true;
        for (int q = 0; q < ns && self_ok; ++q) {
            double v2 = sh[3 * q] * sh[3 * q] + sh[3 * q + 1] * sh[3 * q + 1] + sh[3 * q + 2] * sh[3 * q + 2];
            if (v2 > 0 && v2 < d2min) self_ok = false;
        }
        if (!self_ok) {
            scale *= 0.97;
            continue;
        }
        double s[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
        bool ok = true;
        for (int j = 0; j < m && ok; ++j) {
            double d[3];
            for (int k = 0; k < 3; ++k) {
                d[k] = s[k] - fr[3 * j + k];
                d[k] -= std::nearbyint(d[k]);
            }
            double cx = d[0] * cell[0] + d[1] * cell[3] + d[2] * cell[6];
            double cy = d[0] * cell[1] + d[1] * cell[4] + d[2] * cell[7];
            double cz = d[0] * cell[2] + d[1] * cell[5] + d[2] * cell[8];
            for (int q = 0; q < ns; ++q) {
                double vx = cx + sh[3 * q], vy = cy + sh[3 * q + 1], vz = cz + sh[3 * q + 2];
                if (vx * vx + vy * vy + vz * vz < d2min) {
                    ok = false;
                    break;
                }
            }
        }
        if (ok) {
            for (int k = 0; k < 3; ++k) fr[3 * (size_t)m + k] = s[k];
            ++m;
            stall = 0;
        } else if (++stall >= max_tries) {
            scale *= 0.97;            // jammed: relax the minimum distance
            stall = 0;
        }
    }
    for (int i = 0; i < n; ++i)
        for (int x = 0; x < 3; ++x)
            pos[3 * (size_t)i + x] = fr[3 * i] * cell[x] + fr[3 * i + 1] * cell[3 + x] + fr[3 * i + 2] * cell[6 + x];
    return AGPU_OK;
}
