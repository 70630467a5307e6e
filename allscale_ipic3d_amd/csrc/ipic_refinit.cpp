// ipic_refinit.cpp -- the reference's particle initialisation with its own random streams,
// on the host (one-time set-up; the particles are then uploaded to the GPU).
//
// Restates, with the same streams and the same floating-point steps:
//   initCells (cell.h:512-654) with a generic_particle_generator<vector::uniform,
//   vector::normal, ...> (cell.h:260-300, 405-440), i.e. distribution::uniform_pos_normal_speed
//     * vector::uniform (cell.h:138-175): three std::uniform_real_distribution<double> on one
//       std::minstd_rand; value = generate_canonical<double, 53>(g) * (b - a) + a, which for
//       minstd_rand draws two numbers per value (libstdc++ <bits/random.tcc>);
//     * vector::normal (cell.h:223-257): mean + stddev * ziggurat_normal_distribution()
//       (ziggurat_normal_distribution.h:8-127: Marsaglia & Tsang's RNOR, J. Stat. Softw. 5(8),
//       2000, 128 levels, driven by a 64-bit xorshift state seeded 123456789 ^ seed);
//     * generic_particle_generator::seed (cell.h:278-284): species(s), positions
//       s1 = (1023 s) & s, velocities s2 = (1023 s1) & s1, in 32-bit unsigned arithmetic.
//   Phase 1 (cell.h:525-610): num_particles / (100 ncell) particles per pseudo particle, the
//   pseudo particles drawn from the distribution as constructed (positions seeded seed+1,
//   velocities seed+2), redrawn while outside the universe, binned by int64 truncation of
//   (x - origin) / width (cell.h:63-67); the remainder is spread as missing / ncell per cell plus
//   one for the first missing % ncell cells.
//   Phase 2 (cell.h:613-650): cell (x,y,z) re-seeds a copy of the distribution with
//   (uint32)(((x*1023)+y)*1023+z) and draws its particles' velocities from it, and their
//   positions from a vector::uniform over the cell [low, low + width) with that same seed.
//
// Two streams of the reference do not change any particle: the pseudo particles' velocities
// in phase 1 and the positions drawn by the re-seeded copy in phase 2 (replaced by the cell's
// own uniform).  Both come from generators of their own, so they are not drawn here.  Phase 1
// runs its pseudo particles in parallel chunks, each starting from the LCG state advanced by
// a^(6 i) mod m (six draws per pseudo particle); the first redraw, if any, ends the parallel
// part and the rest runs sequentially from the exact state at that point.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <vector>

#include "ipic_gpu.h"

// every product and sum below is rounded on its own, as the reference's code compiles
#pragma clang fp contract(off)

namespace {

thread_local std::string g_ref_error;

int ref_fail(int rc, const std::string& msg) {
    g_ref_error = msg;
    return rc;
}

// std::minstd_rand: x <- 48271 x mod (2^31 - 1); seed s maps to s mod m, 0 to 1.
constexpr uint64_t kA = 48271, kM = 2147483647;

struct Lcg {
    uint64_t x;
    explicit Lcg(uint64_t s) {
        x = s % kM;
        if (x == 0) x = 1;
    }
    uint64_t next() {
        x = (x * kA) % kM;
        return x;
    }
};

uint64_t powmod(uint64_t b, uint64_t e) {
    uint64_t r = 1;
    b %= kM;
    while (e) {
        if (e & 1) r = (r * b) % kM;
        b = (b * b) % kM;
        e >>= 1;
    }
    return r;
}

// generate_canonical<double, 53> over minstd_rand: range r = max - min + 1 = 2^31 - 2 gives
// floor(log2 r) = 30 bits per draw, hence two draws: (d0 + d1 r) / r^2, each d = draw - min.
// r^2 is rounded to double once, as the library's long double running product is.
constexpr double kR = 2147483646.0;
constexpr double kR2 = 2147483646.0 * 2147483646.0;

inline double canonical(Lcg& g) {
    double sum = (double)(g.next() - 1);
    sum += (double)(g.next() - 1) * kR;
    double ret = sum / kR2;
    if (ret >= 1.0) ret = std::nextafter(1.0, 0.0);
    return ret;
}

inline double uniform(Lcg& g, double a, double b) { return canonical(g) * (b - a) + a; }

// The ziggurat tables (zigset of ziggurat_normal_distribution.h): the same for every seed.
struct ZigTables {
    uint64_t kn[128];
    float wn[128], fn[128];
    ZigTables() {
        const double m1 = 2147483648.0;
        double dn = 3.442619855899, tn = dn;
        const double vn = 9.91256303526217e-3;
        const double q = vn / std::exp(-.5 * dn * dn);
        kn[0] = (uint64_t)((dn / q) * m1);
        kn[1] = 0;
        wn[0] = (float)(q / m1);
        wn[127] = (float)(dn / m1);
        fn[0] = 1.0f;
        fn[127] = (float)std::exp(-.5 * dn * dn);
        for (int i = 126; i >= 1; --i) {
            dn = std::sqrt(-2. * std::log(vn / dn + std::exp(-.5 * dn * dn)));
            kn[i + 1] = (uint64_t)((dn / tn) * m1);
            tn = dn;
            fn[i] = (float)std::exp(-.5 * dn * dn);
            wn[i] = (float)(dn / m1);
        }
    }
};

const ZigTables& zig_tables() {
    static const ZigTables t;
    return t;
}

// One ziggurat stream: the 64-bit xorshift of shr3, and nfix as operator() returns it.  The
// integer conversions (int) / (signed) keep the low 32 bits; uni() and the tail use the float
// logarithm (log of a float argument), the wedge test the double exponential.
struct Zig {
    uint64_t jsr;
    const ZigTables& t;
    explicit Zig(uint64_t seed) : jsr(123456789ull ^ seed), t(zig_tables()) {}
    uint64_t shr3() {
        const uint64_t jz = jsr;
        jsr ^= (jsr << 13);
        jsr ^= (jsr >> 17);
        jsr ^= (jsr << 5);
        return jz + jsr;
    }
    float uni() { return (float)(.5 + (double)(int32_t)(uint32_t)shr3() * .2328306e-9); }
    float operator()() {
        const float r = 3.442620f;
        int32_t hz = (int32_t)(uint32_t)shr3();
        uint32_t iz = (uint32_t)hz & 127u;
        if (std::fabs((double)hz) < (double)t.kn[iz]) return (float)hz * t.wn[iz];
        for (;;) {
            if (iz == 0) {
                float x, y;
                for (;;) {
                    x = (float)(-0.2904764 * (double)std::log(uni()));
                    y = -std::log(uni());
                    if (x * x <= y + y) break;
                }
                return (hz <= 0) ? -r - x : r + x;
            }
            const float x = (float)hz * t.wn[iz];
            if (t.fn[iz] + uni() * (t.fn[iz - 1] - t.fn[iz]) < std::exp(-0.5 * (double)x * (double)x)) return x;
            hz = (int32_t)(uint32_t)shr3();
            iz = (uint32_t)hz & 127u;
            if (std::fabs((double)hz) < (double)t.kn[iz]) return (float)hz * t.wn[iz];
        }
    }
};

struct Geometry {
    int64_t n[3];
    double w[3], origin[3], upper[3];
    int64_t ncell;
};

int geometry(const ipic_desc* d, Geometry& g) {
    if (!d) return ref_fail(IPIC_E_ARG, "null desc");
    g.ncell = 1;
    for (int k = 0; k < 3; ++k) {
        if (d->n[k] <= 0) return ref_fail(IPIC_E_ARG, "grid sizes must be positive");
        if (!(d->cell_width[k] > 0.0)) return ref_fail(IPIC_E_ARG, "cell widths must be positive");
        g.n[k] = d->n[k];
        g.w[k] = d->cell_width[k];
        g.origin[k] = d->origin[k];
        // isInsideUniverse (cell.h:75-83): origin <= x < origin + n * width
        g.upper[k] = d->origin[k] + (double)d->n[k] * d->cell_width[k];
        g.ncell *= g.n[k];
    }
    return 0;
}

int check_dist(const ipic_ref_distribution* dist) {
    if (!dist) return ref_fail(IPIC_E_ARG, "null distribution");
    for (int k = 0; k < 3; ++k)
        if (!std::isfinite(dist->pos_min[k]) || !std::isfinite(dist->pos_max[k]) || !std::isfinite(dist->vel_mean[k]) ||
            !std::isfinite(dist->vel_stddev[k]))
            return ref_fail(IPIC_E_ARG, "non-finite distribution parameters");
    return 0;
}

int thread_count(int32_t want, int64_t work) {
    int t = want > 0 ? want : (int)std::max(1u, std::thread::hardware_concurrency());
    t = (int)std::min<int64_t>(t, std::max<int64_t>(1, work));
    return std::min(t, 256);
}

// The pseudo particle's cell, or -1 when it lies outside the universe (redrawn).
inline int64_t pseudo_cell(const Geometry& g, Lcg& s, const ipic_ref_distribution* d) {
    const double x = uniform(s, d->pos_min[0], d->pos_max[0]);
    const double y = uniform(s, d->pos_min[1], d->pos_max[1]);
    const double z = uniform(s, d->pos_min[2], d->pos_max[2]);
    const double p[3] = {x, y, z};
    int64_t c[3];
    for (int k = 0; k < 3; ++k) {
        if (!(g.origin[k] <= p[k] && p[k] < g.upper[k])) return -1;
        c[k] = (int64_t)((p[k] - g.origin[k]) / g.w[k]);
        if (c[k] >= g.n[k]) c[k] = g.n[k] - 1;  // (x - origin) / width can round up to n
    }
    return (c[0] * g.n[1] + c[1]) * g.n[2] + c[2];
}

}  // namespace

extern "C" {

const char* ipic_ref_last_error(void) { return g_ref_error.c_str(); }

static int cell_counts_impl(const ipic_desc* desc, const ipic_ref_distribution* dist, int64_t* counts) {
    Geometry g;
    if (int rc = geometry(desc, g)) return rc;
    if (int rc = check_dist(dist)) return rc;
    if (!counts) return ref_fail(IPIC_E_ARG, "null counts");
    const uint64_t npseudo = (uint64_t)g.ncell * 100u;
    const uint64_t per = dist->num_particles / npseudo;
    const uint32_t pos_seed = dist->seed + 1u;  // uniform_pos_normal_speed: posGen seed + 1
    std::vector<uint64_t> hist((size_t)g.ncell, 0);
    // parallel part: chunks of pseudo particles from jumped-ahead LCG states
    // one histogram of 4-byte counts per thread: at most 1 GiB of them in all
    const int T = std::max<int>(1, std::min<int64_t>(thread_count(dist->nthreads, (int64_t)(npseudo / 4096 + 1)),
                                                     (int64_t)(1 << 28) / g.ncell));
    std::vector<std::vector<uint32_t>> th(T);
    std::vector<uint64_t> first_redraw(T, npseudo);
    std::vector<std::thread> pool;
    const uint64_t chunk = (npseudo + T - 1) / T;
    const uint64_t x0 = Lcg(pos_seed).x;
    for (int t = 0; t < T; ++t) {
        pool.emplace_back([&, t]() {
            const uint64_t i0 = std::min<uint64_t>(npseudo, chunk * t), i1 = std::min<uint64_t>(npseudo, i0 + chunk);
            if (i0 >= i1) return;
            th[t].assign((size_t)g.ncell, 0u);
            Lcg s(1);
            s.x = (x0 * powmod(kA, 6 * i0)) % kM;
            for (uint64_t i = i0; i < i1; ++i) {
                const int64_t c = pseudo_cell(g, s, dist);
                if (c < 0) {
                    first_redraw[t] = i;
                    return;
                }
                th[t][(size_t)c]++;
            }
        });
    }
    for (auto& p : pool) p.join();
    uint64_t done = npseudo;
    for (int t = 0; t < T; ++t) {
        if (!th[t].empty())
            for (int64_t c = 0; c < g.ncell; ++c) hist[(size_t)c] += th[t][(size_t)c];
        if (first_redraw[t] < npseudo) {
            done = first_redraw[t];  // chunks after this one started from the wrong state
            break;
        }
    }
    if (done < npseudo) {
        // sequential from the first redraw: the exact state after 6 * done draws
        Lcg s(1);
        s.x = (x0 * powmod(kA, 6 * done)) % kM;
        for (uint64_t i = done; i < npseudo; ++i) {
            int64_t c;
            uint64_t tries = 0;
            while ((c = pseudo_cell(g, s, dist)) < 0)
                if (++tries > 100000000ull) return ref_fail(IPIC_E_ARG, "the position box does not meet the universe");
            hist[(size_t)c]++;
        }
    }
    uint64_t sum = 0;
    for (int64_t c = 0; c < g.ncell; ++c) {
        counts[c] = (int64_t)(hist[(size_t)c] * per);
        sum += hist[(size_t)c] * per;
    }
    const int64_t missing = (int64_t)(dist->num_particles - sum);
    // the reference asserts missing > 0 (cell.h:557)
    if (missing <= 0) return ref_fail(IPIC_E_ARG, "initCells needs num_particles not a multiple of 100 * ncell per pseudo particle");
    const int64_t all = missing / g.ncell, rem = missing % g.ncell;
    for (int64_t c = 0; c < g.ncell; ++c) counts[c] += all + (c < rem ? 1 : 0);
    return 0;
}

static int particles_impl(const ipic_desc* desc, const ipic_ref_distribution* dist, const int64_t* counts,
                          double* pos, double* vel) {
    Geometry g;
    if (int rc = geometry(desc, g)) return rc;
    if (int rc = check_dist(dist)) return rc;
    if (!counts || !pos || !vel) return ref_fail(IPIC_E_ARG, "null counts or output arrays");
    std::vector<int64_t> start((size_t)g.ncell + 1, 0);
    for (int64_t c = 0; c < g.ncell; ++c) {
        if (counts[c] < 0) return ref_fail(IPIC_E_ARG, "negative cell count");
        start[(size_t)c + 1] = start[(size_t)c] + counts[c];
    }
    const int T = thread_count(dist->nthreads, g.ncell);
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t) {
        pool.emplace_back([&]() {
            for (;;) {
                const int64_t c0 = next.fetch_add(64);
                if (c0 >= g.ncell) return;
                const int64_t c1 = std::min<int64_t>(g.ncell, c0 + 64);
                for (int64_t c = c0; c < c1; ++c) {
                    const int64_t x = c / (g.n[1] * g.n[2]), y = (c / g.n[2]) % g.n[1], z = c % g.n[2];
                    const uint32_t seed = (uint32_t)(((x * 1023) + y) * 1023 + z);
                    const uint32_t s1 = (1023u * seed) & seed;  // the copy's position stream (unused)
                    const uint32_t s2 = (1023u * s1) & s1;      // the copy's velocity stream
                    Zig vz(s2);
                    Lcg ps(seed);
                    const double low[3] = {g.w[0] * (double)x + g.origin[0], g.w[1] * (double)y + g.origin[1],
                                           g.w[2] * (double)z + g.origin[2]};
                    const double hig[3] = {low[0] + g.w[0], low[1] + g.w[1], low[2] + g.w[2]};
                    for (int64_t i = start[(size_t)c]; i < start[(size_t)c + 1]; ++i) {
                        for (int k = 0; k < 3; ++k)
                            vel[3 * i + k] = dist->vel_mean[k] + dist->vel_stddev[k] * (double)vz();
                        for (int k = 0; k < 3; ++k) pos[3 * i + k] = uniform(ps, low[k], hig[k]);
                    }
                }
            }
        });
    }
    for (auto& p : pool) p.join();
    return 0;
}

// Parameter-file initCells (cell.h:699-776).  Every cell draws from its own minstd stream, so
// cells run on any thread in any order; the k-th particle of cell c lands at c * per + k.
static int lattice_impl(const ipic_desc* desc, const ipic_ref_lattice* lat, double* pos, double* vel) {
    Geometry g;
    if (int rc = geometry(desc, g)) return rc;
    if (!lat) return ref_fail(IPIC_E_ARG, "null lattice description");
    if (!pos || !vel) return ref_fail(IPIC_E_ARG, "null output arrays");
    for (int k = 0; k < 3; ++k) {
        if (lat->ppc[k] <= 0) return ref_fail(IPIC_E_ARG, "particles per cell must be positive on every axis");
        if (!std::isfinite(lat->vel_thermal[k]) || !std::isfinite(lat->vel_drift[k]))
            return ref_fail(IPIC_E_ARG, "non-finite thermal or drift velocity");
    }
    const int64_t per = (int64_t)lat->ppc[0] * lat->ppc[1] * lat->ppc[2];
    double sub[3];
    for (int k = 0; k < 3; ++k) sub[k] = g.w[k] / (double)lat->ppc[k];
    const int T = thread_count(lat->nthreads, g.ncell);
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t) {
        pool.emplace_back([&]() {
            for (;;) {
                const int64_t c0 = next.fetch_add(64);
                if (c0 >= g.ncell) return;
                for (int64_t c = c0; c < std::min<int64_t>(g.ncell, c0 + 64); ++c) {
                    const int64_t cx = c / (g.n[1] * g.n[2]), cy = (c / g.n[2]) % g.n[1], cz = c % g.n[2];
                    const double corner[3] = {g.origin[0] + (double)cx * g.w[0], g.origin[1] + (double)cy * g.w[1],
                                              g.origin[2] + (double)cz * g.w[2]};
                    // the seed is the cell's position folded as x * 10000 + y * 100 + z (unsigned)
                    Lcg s((uint32_t)(cx * 10000 + cy * 100 + cz));
                    auto unit = [&s]() { return (double)s.next() / (double)(kM - 1); };
                    int64_t i = c * per;
                    for (int32_t a = 0; a < lat->ppc[0]; ++a)
                        for (int32_t b = 0; b < lat->ppc[1]; ++b)
                            for (int32_t d = 0; d < lat->ppc[2]; ++d, ++i) {
                                const int32_t l[3] = {a, b, d};
                                for (int k = 0; k < 3; ++k) pos[3 * i + k] = ((double)l[k] + 0.5) * sub[k] + corner[k];
                                // Box-Muller: two radii and two angles per particle, drawn r, phi, r, phi
                                const double r0 = std::sqrt(-2.0 * std::log(1.0 - 0.999999 * unit()));
                                const double p0 = 2.0 * M_PI * unit();
                                const double r1 = std::sqrt(-2.0 * std::log(1.0 - 0.999999 * unit()));
                                const double p1 = 2.0 * M_PI * unit();
                                vel[3 * i + 0] = lat->vel_drift[0] + lat->vel_thermal[0] * (r0 * std::cos(p0));
                                vel[3 * i + 1] = lat->vel_drift[1] + lat->vel_thermal[1] * (r0 * std::sin(p0));
                                vel[3 * i + 2] = lat->vel_drift[2] + lat->vel_thermal[2] * (r1 * std::cos(p1));
                            }
                }
            }
        });
    }
    for (auto& p : pool) p.join();
    return 0;
}

// no C++ exception crosses the C ABI: a failed allocation or thread start is an error code
int ipic_ref_lattice_particles(const ipic_desc* desc, const ipic_ref_lattice* lat, double* pos, double* vel) {
    try {
        return lattice_impl(desc, lat, pos, vel);
    } catch (const std::exception& e) {
        return ref_fail(IPIC_E_RANGE, std::string("reference init: ") + e.what());
    }
}

double ipic_ref_lattice_charge(const double cell_width[3], const int32_t ppc[3], double qom, double rho) {
    // q_factor of cell.h:711-716, the same operations in the same order
    const double sign = qom / std::fabs(qom);
    const double vol = cell_width[0] * cell_width[1] * cell_width[2];
    return sign * vol / (double)((int64_t)ppc[0] * ppc[1] * ppc[2]) * rho / (16.0 * std::atan(1.0));
}

int ipic_ref_cell_counts(const ipic_desc* desc, const ipic_ref_distribution* dist, int64_t* counts) {
    try {
        return cell_counts_impl(desc, dist, counts);
    } catch (const std::exception& e) {
        return ref_fail(IPIC_E_RANGE, std::string("reference init: ") + e.what());
    }
}

int ipic_ref_particles(const ipic_desc* desc, const ipic_ref_distribution* dist, const int64_t* counts, double* pos,
                       double* vel) {
    try {
        return particles_impl(desc, dist, counts, pos, vel);
    } catch (const std::exception& e) {
        return ref_fail(IPIC_E_RANGE, std::string("reference init: ") + e.what());
    }
}

}  // extern "C"
