// initconf.cpp -- seeded initial state (host only).
//
// The reference only has lattice starts (init-cube1.f:33-128) seeded from the
// time of day or a fixed constant (kps.f:22-30).  The nucleation workloads
// need seeded random vapour instead: uniform positions with a hard minimum
// distance, Maxwell-Boltzmann velocities, zero momentum, and kinetic energy
// exactly 3/2 N k T (the target of thermostat(-1), scalev-m.f / stopdead-m.f).
//
// Random numbers: xoshiro256** (Blackman & Vigna 2018) seeded through
// SplitMix64; normals by the Box-Muller transform.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/fcmd.h"

namespace {

struct Xoshiro {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    explicit Xoshiro(uint64_t seed)
    {
        for (auto &v : s) v = splitmix(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
    double normal()
    {
        double u1;
        do u1 = uniform();
        while (u1 <= 0.0);
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

constexpr double CIKB = 1.38066e-6;  // const.inc:17

}  // namespace

extern "C" int fcmd_init_random(unsigned long long seed, int natom, const double *box, double rmin,
                                double att, const double *mas, const int *api, double *x, double *y,
                                double *z, double *vx, double *vy, double *vz)
{
    if (natom <= 0 || !box || !mas || !api || !x || !y || !z || !vx || !vy || !vz || !(rmin >= 0))
        return FCMD_ERR_ARG;
    Xoshiro rng(seed);
    const double r2min = rmin * rmin;
    int nc[3];
    double side[3];
    for (int d = 0; d < 3; ++d) {
        if (!(box[d] > 0)) return FCMD_ERR_ARG;
        nc[d] = rmin > 0 ? (int)std::floor(box[d] / rmin) : 1;
        if (nc[d] < 1) nc[d] = 1;
        if (nc[d] > 1024) nc[d] = 1024;
        side[d] = box[d] / nc[d];
    }
    const long long ncell = (long long)nc[0] * nc[1] * nc[2];
    std::vector<int> head(ncell, -1), next(natom, -1);
    auto cell_of = [&](double p, int d) {
        int c = (int)(p / side[d]);
        return c < 0 ? 0 : (c >= nc[d] ? nc[d] - 1 : c);
    };
    const long long max_tries = 2000LL * natom + 100000;
    long long tries = 0;
    for (int i = 0; i < natom;) {
        if (++tries > max_tries) return FCMD_ERR_ARG;
        const double p[3] = {rng.uniform() * box[0], rng.uniform() * box[1], rng.uniform() * box[2]};
        const int c[3] = {cell_of(p[0], 0), cell_of(p[1], 1), cell_of(p[2], 2)};
        bool ok = true;
        for (int oz = -1; oz <= 1 && ok; ++oz)
            for (int oy = -1; oy <= 1 && ok; ++oy)
                for (int ox = -1; ox <= 1 && ok; ++ox) {
                    // stencil without duplicates when a dimension has < 3 cells
                    if ((nc[0] < 3 && ox != 0 && (nc[0] == 1 || ox == -1)) ||
                        (nc[1] < 3 && oy != 0 && (nc[1] == 1 || oy == -1)) ||
                        (nc[2] < 3 && oz != 0 && (nc[2] == 1 || oz == -1)))
                        continue;
                    const int cx = (c[0] + ox + nc[0]) % nc[0];
                    const int cy = (c[1] + oy + nc[1]) % nc[1];
                    const int cz = (c[2] + oz + nc[2]) % nc[2];
                    for (int j = head[((long long)cz * nc[1] + cy) * nc[0] + cx]; j >= 0; j = next[j]) {
                        double d[3] = {p[0] - x[j], p[1] - y[j], p[2] - z[j]};
                        for (int k = 0; k < 3; ++k) d[k] -= box[k] * std::nearbyint(d[k] / box[k]);
                        if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < r2min) {
                            ok = false;
                            break;
                        }
                    }
                }
        if (!ok) continue;
        x[i] = p[0];
        y[i] = p[1];
        z[i] = p[2];
        const long long cell = ((long long)c[2] * nc[1] + c[1]) * nc[0] + c[0];
        next[i] = head[cell];
        head[cell] = i;
        ++i;
    }
    // Maxwell-Boltzmann velocities, zero momentum, exact temperature
    double px = 0, py = 0, pz = 0, mt = 0;
    for (int i = 0; i < natom; ++i) {
        const double m = mas[api[i] - 1];
        const double sd = std::sqrt(CIKB * att / m);
        vx[i] = rng.normal() * sd;
        vy[i] = rng.normal() * sd;
        vz[i] = rng.normal() * sd;
        px += m * vx[i];
        py += m * vy[i];
        pz += m * vz[i];
        mt += m;
    }
    double ek = 0;
    for (int i = 0; i < natom; ++i) {
        const double m = mas[api[i] - 1];
        vx[i] -= px / mt;
        vy[i] -= py / mt;
        vz[i] -= pz / mt;
        ek += 0.5 * m * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    }
    const double target = 1.5 * CIKB * att * natom;
    const double f = ek > 0 ? std::sqrt(target / ek) : 0.0;
    for (int i = 0; i < natom; ++i) {
        vx[i] *= f;
        vy[i] *= f;
        vz[i] *= f;
    }
    return FCMD_OK;
}
