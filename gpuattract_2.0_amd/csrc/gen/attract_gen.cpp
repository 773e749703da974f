// Seeded synthetic-workload generator (host C++).  See include/attract_gen.h
// for the contract.  Deterministic: every random draw is a pure function of
// (seed, entity, counter) and every floating-point sum runs in a fixed order,
// so results do not depend on the number of threads.
#include "attract_gen.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t draw(uint64_t seed, unsigned entity, uint64_t counter) {
    uint64_t key = mix64(seed ^ mix64(0xA5A5A5A5ull + entity));
    return mix64(key + counter * 0xD1B54A32D192ED03ull);
}

inline double unif(uint64_t seed, unsigned entity, uint64_t counter) {
    return (double)(draw(seed, entity, counter) >> 11) * (1.0 / 9007199254740992.0);
}

// Standard normal from two uniforms at counters 2c, 2c+1 (Box-Muller, cosine branch).
inline double gauss(uint64_t seed, unsigned entity, uint64_t c) {
    double u1 = unif(seed, entity, 2 * c);
    double u2 = unif(seed, entity, 2 * c + 1);
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
}

// Parallel for over [0, n) in contiguous chunks; each index is handled by
// exactly one thread, so per-index results do not depend on nthreads.
template <class F>
void parallel_for(size_t n, int nthreads, F&& f) {
    if (nthreads <= 1 || n < 2) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    size_t chunk = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        size_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([lo, hi, &f] {
            for (size_t i = lo; i < hi; ++i) f(i);
        });
    }
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" double attg_uniform(uint64_t seed, unsigned entity, uint64_t counter) {
    return unif(seed, entity, counter);
}

extern "C" int attg_make_body(uint64_t seed, unsigned entity, unsigned n, double radius,
                              double min_spacing, unsigned num_types, double charged_frac,
                              double type0_frac, float* xyz, unsigned* type, float* charge) {
    // Sub-streams: positions, types, charges, type-0 draws.
    const unsigned ePos = entity * 16 + 0, eTyp = entity * 16 + 1, eChg = entity * 16 + 2,
                   eZero = entity * 16 + 3;
    std::vector<double> px, py, pz;
    px.reserve(n); py.reserve(n); pz.reserve(n);
    // Cell list with cell edge = min_spacing over the cube [-R, R]^3.
    const double cell = std::max(min_spacing, 1e-3);
    const int nc = std::max(1, (int)std::ceil(2.0 * radius / cell));
    std::vector<std::vector<unsigned>> cells((size_t)nc * nc * nc);
    auto cidx = [&](double v) {
        int c = (int)std::floor((v + radius) / cell);
        return std::min(std::max(c, 0), nc - 1);
    };
    const double ms2 = min_spacing * min_spacing;
    uint64_t counter = 0;
    const uint64_t maxTries = (uint64_t)n * 20000ull + 100000ull;
    while (px.size() < n) {
        if (counter / 3 > maxTries) return -1;
        double x = (2.0 * unif(seed, ePos, counter++) - 1.0) * radius;
        double y = (2.0 * unif(seed, ePos, counter++) - 1.0) * radius;
        double z = (2.0 * unif(seed, ePos, counter++) - 1.0) * radius;
        if (x * x + y * y + z * z > radius * radius) continue;
        int cx = cidx(x), cy = cidx(y), cz = cidx(z);
        bool ok = true;
        for (int dz = -1; dz <= 1 && ok; ++dz)
            for (int dy = -1; dy <= 1 && ok; ++dy)
                for (int dx = -1; dx <= 1 && ok; ++dx) {
                    int ax = cx + dx, ay = cy + dy, az = cz + dz;
                    if (ax < 0 || ay < 0 || az < 0 || ax >= nc || ay >= nc || az >= nc) continue;
                    for (unsigned k : cells[((size_t)az * nc + ay) * nc + ax]) {
                        double ddx = x - px[k], ddy = y - py[k], ddz = z - pz[k];
                        if (ddx * ddx + ddy * ddy + ddz * ddz < ms2) { ok = false; break; }
                    }
                }
        if (!ok) continue;
        cells[((size_t)cz * nc + cy) * nc + cx].push_back((unsigned)px.size());
        px.push_back(x); py.push_back(y); pz.push_back(z);
    }
    double sx = 0, sy = 0, sz = 0;
    for (unsigned i = 0; i < n; ++i) { sx += px[i]; sy += py[i]; sz += pz[i]; }
    sx /= n; sy /= n; sz /= n;
    for (unsigned i = 0; i < n; ++i) {
        xyz[i] = (float)(px[i] - sx);
        xyz[n + i] = (float)(py[i] - sy);
        xyz[2 * n + i] = (float)(pz[i] - sz);
        unsigned t = 1 + (unsigned)(unif(seed, eTyp, i) * num_types);
        if (t > num_types) t = num_types;
        if (type0_frac > 0 && unif(seed, eZero, i) < type0_frac) t = 0;
        type[i] = t;
        double u = unif(seed, eChg, 2 * (uint64_t)i);
        float q = 0.f;
        if (u < charged_frac) q = unif(seed, eChg, 2 * (uint64_t)i + 1) < 0.5 ? -1.f : 1.f;
        charge[i] = q;
    }
    return 0;
}

extern "C" void attg_make_raw_params(uint64_t seed, unsigned T, double rmin_lo, double rmin_hi,
                                     double emin_lo, double emin_hi, double repulsive_frac,
                                     float* rbc, float* abc, int* ipon) {
    const unsigned e = ATTG_ENT_TABLE;
    for (unsigned a = 0; a < T; ++a)
        for (unsigned b = a; b < T; ++b) {
            uint64_t c = ((uint64_t)a * T + b) * 4;
            double rmin = rmin_lo + (rmin_hi - rmin_lo) * unif(seed, e, c);
            double emin = emin_lo + (emin_hi - emin_lo) * unif(seed, e, c + 1);
            int ip = unif(seed, e, c + 2) < repulsive_frac ? -1 : 1;
            // Inverse of the 8-6 derivation: rmin^2 = 4/3 rbc^2, |emin| = 27/256 abc.
            float r = (float)(rmin * std::sqrt(0.75));
            float A = (float)(emin * 256.0 / 27.0);
            rbc[a * T + b] = rbc[b * T + a] = r;
            abc[a * T + b] = abc[b * T + a] = A;
            ipon[a * T + b] = ipon[b * T + a] = ip;
        }
}

// fileIO/readFile.cpp:479-513: rc = abc*rbc^shape, ac = abc*rbc^6 (std::pow with an
// unsigned exponent promotes to double), then emin/rmin2 from rc, ac.
extern "C" int attg_derive_params(unsigned T, int shape, const float* rbc, const float* abc,
                                  const int* ipon_in, float* rc, float* ac, int* ipon,
                                  float* rmin2, float* emin) {
    if (shape != 8 && shape != 12) return -1;
    for (unsigned k = 0; k < T * T; ++k) {
        const float r = rbc[k], A = abc[k];
        ipon[k] = ipon_in[k];
        if (shape == 8) {
            float rck = (float)(A * std::pow((double)r, 8.0));
            float ack = (float)(A * std::pow((double)r, 6.0));
            rc[k] = rck;
            ac[k] = ack;
            if (ack > 0 && rck > 0) {
                emin[k] = (float)(-27.0f * std::pow((double)ack, 4.0) /
                                  (256.0f * std::pow((double)rck, 3.0)));
                rmin2[k] = 4.0f * rck / (3.0f * ack);
            } else {
                emin[k] = 0;
                rmin2[k] = 0;
            }
        } else {
            rc[k] = (float)(A * std::pow((double)r, 12.0));
            ac[k] = (float)(A * std::pow((double)r, 6.0));
            emin[k] = (float)(-0.25 * A);
            rmin2[k] = 1.25992105f * r * r;
        }
    }
    return 0;
}

extern "C" void attg_inner_geom(const float* xyz, unsigned n, float spacing, float margin,
                                attg_geom* g) {
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = 1e30; hi[d] = -1e30;
        for (unsigned i = 0; i < n; ++i) {
            lo[d] = std::min(lo[d], (double)xyz[d * n + i]);
            hi[d] = std::max(hi[d], (double)xyz[d * n + i]);
        }
    }
    unsigned* dims[3] = {&g->nx, &g->ny, &g->nz};
    float* org[3] = {&g->ox, &g->oy, &g->oz};
    for (int d = 0; d < 3; ++d) {
        double ext = hi[d] - lo[d] + 2.0 * margin;
        *dims[d] = (unsigned)std::ceil(ext / spacing) + 1;
        *org[d] = (float)(lo[d] - margin);
    }
    g->spacing = spacing;
}

extern "C" void attg_cube_geom(float cx, float cy, float cz, float halfwidth, float spacing,
                               attg_geom* g) {
    unsigned nd = (unsigned)std::ceil(2.0 * halfwidth / spacing) + 1;
    g->nx = g->ny = g->nz = nd;
    double half = 0.5 * (nd - 1) * spacing;
    g->ox = (float)(cx - half);
    g->oy = (float)(cy - half);
    g->oz = (float)(cz - half);
    g->spacing = spacing;
}

namespace {

struct PairP {
    double rc, ac, ipon, rmin2, emin;
};

// Pair potential at squared distance dr2 (already max'ed with the plateau), in
// double: the semantics of LJPotForce (model/forcefield.h:28-83).  Returns the
// energy and fb, the factor with force = fb * d / (r * r').
inline void lj_eval(double dr2, const PairP& p, int shape, double& E, double& fb) {
    double dr2_inv = 1.0 / dr2;
    double rr23 = dr2_inv * dr2_inv * dr2_inv;
    double rrd = shape == 8 ? dr2_inv : rr23;
    double sd = shape == 8 ? 2.0 : 6.0;
    double rep = p.rc * rrd;
    double vlj = (rep - p.ac) * rr23;
    double f = 6.0 * vlj + sd * (rep * rr23);
    if (dr2 < p.rmin2) {
        E = vlj + (p.ipon - 1.0) * p.emin;
        fb = f;
    } else {
        E = p.ipon * vlj;
        fb = p.ipon * f;
    }
}

}  // namespace


extern "C" int attg_build_grid(const attg_geom* g, const float* xyz, const unsigned* type,
                               const float* charge, unsigned n, unsigned T, int shape,
                               const float* rc, const float* ac, const int* ipon,
                               const float* rmin2, const float* emin, int src_is_ligand,
                               float plateau, int dielec, float ffelec, float* out,
                               int nthreads) {
    if (shape != 8 && shape != 12) return -1;
    const size_t nx = g->nx, ny = g->ny, nz = g->nz, npts = nx * ny * nz;
    const unsigned TT = T + 1;
    // Pair parameters seen from (source type a, probe type t).
    std::vector<PairP> tab((size_t)TT * TT);
    const double P2 = (double)plateau * (double)plateau;
    bool linear = true;  // every pair stays in the r >= rmin branch once r' >= plateau
    for (unsigned a = 1; a <= T; ++a)
        for (unsigned t = 1; t <= T; ++t) {
            unsigned k = src_is_ligand ? (t - 1) * T + (a - 1) : (a - 1) * T + (t - 1);
            tab[a * TT + t] = PairP{rc[k], ac[k], (double)ipon[k], rmin2[k], emin[k]};
            if ((double)rmin2[k] > P2) linear = false;
        }
    // Source beads grouped for the inner loops.
    std::vector<unsigned> act, chg;
    for (unsigned j = 0; j < n; ++j) {
        if (type[j] == 0 || type[j] > T) continue;
        act.push_back(j);
        if (charge[j] != 0.f) chg.push_back(j);
    }
    const float* X = xyz;
    const float* Y = xyz + n;
    const float* Z = xyz + 2 * (size_t)n;
    const double sd = shape == 8 ? 2.0 : 6.0;
    const double fe = ffelec;

    parallel_for(npts, nthreads, [&](size_t pidx) {
        const size_t ix = pidx % nx, iy = (pidx / nx) % ny, iz = pidx / (nx * ny);
        const double px = (double)g->ox + (double)ix * g->spacing;
        const double py = (double)g->oy + (double)iy * g->spacing;
        const double pz = (double)g->oz + (double)iz * g->spacing;
        // Per source type: sum of 1/r'^s (s = 8 or 12), 1/r'^6 and the matching
        // force directions d/(r r' r'^s), d/(r r' r'^6).
        thread_local std::vector<double> accV, EtV, FtV;
        accV.assign((size_t)TT * 8, 0.0);
        EtV.assign(TT, 0.0);
        FtV.assign((size_t)TT * 3, 0.0);
        double* acc = accV.data();
        double* Et = EtV.data();
        double (*Ft)[3] = reinterpret_cast<double (*)[3]>(FtV.data());
        if (linear) {
            std::fill(acc, acc + TT * 8, 0.0);
            for (unsigned j : act) {
                const double dx = px - X[j], dy = py - Y[j], dz = pz - Z[j];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double r = std::sqrt(r2);
                const double rp2 = r2 > P2 ? r2 : P2;
                const double rp = r2 > P2 ? r : std::sqrt(P2);
                const double i2 = 1.0 / rp2;
                const double i6 = i2 * i2 * i2;
                const double is = shape == 8 ? i6 * i2 : i6 * i6;
                double* A = acc + type[j] * 8;
                A[0] += is;
                A[1] += i6;
                if (r > 0) {
                    const double w = 1.0 / (r * rp);
                    A[2] += is * w * dx; A[3] += is * w * dy; A[4] += is * w * dz;
                    A[5] += i6 * w * dx; A[6] += i6 * w * dy; A[7] += i6 * w * dz;
                }
            }
            for (unsigned t = 1; t <= T; ++t)
                for (unsigned a = 1; a <= T; ++a) {
                    const PairP& p = tab[a * TT + t];
                    const double* A = acc + a * 8;
                    Et[t] += p.ipon * (p.rc * A[0] - p.ac * A[1]);
                    for (int c = 0; c < 3; ++c)
                        Ft[t][c] += p.ipon * ((6.0 + sd) * p.rc * A[2 + c] - 6.0 * p.ac * A[5 + c]);
                }
        } else {
            for (unsigned j : act) {
                const double dx = px - X[j], dy = py - Y[j], dz = pz - Z[j];
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double r = std::sqrt(r2);
                const double rp2 = r2 > P2 ? r2 : P2;
                const double rp = r2 > P2 ? r : std::sqrt(P2);
                const double w = r > 0 ? 1.0 / (r * rp) : 0.0;
                for (unsigned t = 1; t <= T; ++t) {
                    double E, fb;
                    lj_eval(rp2, tab[type[j] * TT + t], shape, E, fb);
                    Et[t] += E;
                    Ft[t][0] += fb * w * dx; Ft[t][1] += fb * w * dy; Ft[t][2] += fb * w * dz;
                }
            }
        }
        // Electrostatics (grid 0), model/forcefield.h:85-146 with r' = max(r, plateau).
        for (unsigned j : chg) {
            const double dx = px - X[j], dy = py - Y[j], dz = pz - Z[j];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double r = std::sqrt(r2);
            const double rp2 = r2 > P2 ? r2 : P2;
            const double rp = r2 > P2 ? r : std::sqrt(P2);
            const double c = (double)charge[j] * fe;
            double dd, fmag;  // energy factor and |F| * r'
            if (dielec == 0) {
                dd = 1.0 / rp - 1.0 / 50.0;
                fmag = c / rp;
            } else {
                dd = 1.0 / rp2 - 1.0 / 2500.0;
                fmag = 2.0 * c / rp2;
            }
            if (dd <= 0) continue;
            Et[0] += c * dd;
            if (r > 0) {
                const double w = fmag / (r * rp);
                Ft[0][0] += w * dx; Ft[0][1] += w * dy; Ft[0][2] += w * dz;
            }
        }
        for (unsigned t = 0; t < TT; ++t) {
            float* o = out + ((size_t)t * npts + pidx) * 4;
            o[0] = (float)Ft[t][0];
            o[1] = (float)Ft[t][1];
            o[2] = (float)Ft[t][2];
            o[3] = (float)Et[t];
        }
    });
    return 0;
}

extern "C" int attg_build_basis(const attg_geom* g, const float* xyz, const unsigned* type,
                                unsigned n, unsigned T, int shape, const float* rmin2,
                                float plateau, float* out, int nthreads) {
    if (shape != 8 && shape != 12) return -2;
    const double P2 = (double)plateau * (double)plateau;
    for (unsigned k = 0; k < T * T; ++k)
        if ((double)rmin2[k] > P2) return -1;
    const size_t nx = g->nx, ny = g->ny, nz = g->nz, npts = nx * ny * nz;
    const float* X = xyz;
    const float* Y = xyz + n;
    const float* Z = xyz + 2 * (size_t)n;
    parallel_for(npts, nthreads, [&](size_t pidx) {
        const size_t ix = pidx % nx, iy = (pidx / nx) % ny, iz = pidx / (nx * ny);
        const double px = (double)g->ox + (double)ix * g->spacing;
        const double py = (double)g->oy + (double)iy * g->spacing;
        const double pz = (double)g->oz + (double)iz * g->spacing;
        double acc[2 * 100];
        std::fill(acc, acc + 2 * T, 0.0);
        for (unsigned j = 0; j < n; ++j) {
            const unsigned a = type[j];
            if (a == 0 || a > T) continue;
            const double dx = px - X[j], dy = py - Y[j], dz = pz - Z[j];
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double i2 = 1.0 / (r2 > P2 ? r2 : P2);
            const double i6 = i2 * i2 * i2;
            acc[2 * (a - 1)] += shape == 8 ? i6 * i2 : i6 * i6;
            acc[2 * (a - 1) + 1] += i6;
        }
        for (unsigned b = 0; b < 2 * T; ++b) out[b * npts + pidx] = (float)acc[b];
    });
    return 0;
}

extern "C" int attg_build_nl(const attg_geom* g, const float* xyz, const unsigned* type,
                             unsigned n, float neighbourdis, uint32_t* desc, uint32_t* list,
                             uint64_t* total, int nthreads) {
    const size_t nx = g->nx, ny = g->ny, nz = g->nz, npts = nx * ny * nz;
    const double nd = neighbourdis, nd2 = nd * nd;
    const float* X = xyz;
    const float* Y = xyz + n;
    const float* Z = xyz + 2 * (size_t)n;
    // Cell list of the source beads, cell edge = neighbourdis.
    double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
    for (unsigned j = 0; j < n; ++j) {
        const double v[3] = {X[j], Y[j], Z[j]};
        for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], v[d]); hi[d] = std::max(hi[d], v[d]); }
    }
    int nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = std::max(1, (int)std::ceil((hi[d] - lo[d]) / nd) + 1);
    std::vector<std::vector<unsigned>> cells((size_t)nc[0] * nc[1] * nc[2]);
    for (unsigned j = 0; j < n; ++j) {
        if (type[j] == 0) continue;
        int c[3] = {(int)((X[j] - lo[0]) / nd), (int)((Y[j] - lo[1]) / nd), (int)((Z[j] - lo[2]) / nd)};
        cells[((size_t)c[2] * nc[1] + c[1]) * nc[0] + c[0]].push_back(j);
    }
    auto collect = [&](size_t pidx, std::vector<unsigned>& found) {
        found.clear();
        const size_t ix = pidx % nx, iy = (pidx / nx) % ny, iz = pidx / (nx * ny);
        const double p[3] = {(double)g->ox + (double)ix * g->spacing,
                             (double)g->oy + (double)iy * g->spacing,
                             (double)g->oz + (double)iz * g->spacing};
        int c0[3], c1[3];
        for (int d = 0; d < 3; ++d) {
            c0[d] = std::max(0, (int)std::floor((p[d] - nd - lo[d]) / nd));
            c1[d] = std::min(nc[d] - 1, (int)std::floor((p[d] + nd - lo[d]) / nd));
            if (c1[d] < c0[d]) return;
        }
        for (int cz = c0[2]; cz <= c1[2]; ++cz)
            for (int cy = c0[1]; cy <= c1[1]; ++cy)
                for (int cx = c0[0]; cx <= c1[0]; ++cx)
                    for (unsigned j : cells[((size_t)cz * nc[1] + cy) * nc[0] + cx]) {
                        const double dx = p[0] - X[j], dy = p[1] - Y[j], dz = p[2] - Z[j];
                        if (dx * dx + dy * dy + dz * dz <= nd2) found.push_back(j);
                    }
        std::sort(found.begin(), found.end());
    };
    // Pass 1: counts.
    std::vector<uint32_t> cnt(npts);
    parallel_for(npts, nthreads, [&](size_t p) {
        thread_local std::vector<unsigned> f;
        collect(p, f);
        cnt[p] = (uint32_t)f.size();
    });
    uint64_t off = 0;
    for (size_t p = 0; p < npts; ++p) {
        desc[2 * p] = cnt[p];
        desc[2 * p + 1] = (uint32_t)off;
        off += cnt[p];
    }
    if (total) *total = off;
    if (off > 0xFFFFFFFFull) return -1;
    if (list == nullptr) return 0;
    parallel_for(npts, nthreads, [&](size_t p) {
        thread_local std::vector<unsigned> f;
        collect(p, f);
        std::copy(f.begin(), f.end(), list + desc[2 * p + 1]);
    });
    return 0;
}

extern "C" void attg_make_poses(uint64_t seed, unsigned entity, uint64_t first, uint64_t n,
                                double r_lo, double r_hi, float cx, float cy, float cz,
                                float* dof) {
    for (uint64_t k = 0; k < n; ++k) {
        const uint64_t c = (first + k) * 8;
        const double phi = kTwoPi * unif(seed, entity, c + 0);
        const double rot = kTwoPi * unif(seed, entity, c + 1);
        const double ssi = std::acos(2.0 * unif(seed, entity, c + 2) - 1.0);
        const double uz = 2.0 * unif(seed, entity, c + 3) - 1.0;
        const double az = kTwoPi * unif(seed, entity, c + 4);
        const double rad = r_lo + (r_hi - r_lo) * unif(seed, entity, c + 5);
        const double s = std::sqrt(std::max(0.0, 1.0 - uz * uz));
        float* d = dof + 6 * k;
        d[0] = (float)(cx + rad * s * std::cos(az));
        d[1] = (float)(cy + rad * s * std::sin(az));
        d[2] = (float)(cz + rad * uz);
        d[3] = (float)phi;
        d[4] = (float)ssi;
        d[5] = (float)rot;
    }
}

extern "C" void attg_make_amplitudes(uint64_t seed, unsigned entity, uint64_t first, uint64_t n,
                                     unsigned M, double sigma, float* out) {
    for (uint64_t k = 0; k < n * M; ++k) out[k] = (float)(sigma * gauss(seed, entity, first * M + k));
}

extern "C" void attg_make_modes(uint64_t seed, unsigned entity, unsigned n, unsigned M,
                                double h_lo, double h_hi, float* modes, float* force_const) {
    const size_t len = 3 * (size_t)n;
    std::vector<double> V((size_t)M * len);
    for (unsigned m = 0; m < M; ++m)
        for (size_t c = 0; c < len; ++c) V[m * len + c] = gauss(seed, entity, m * len + c);
    // Modified Gram-Schmidt, two passes.
    for (int pass = 0; pass < 2; ++pass)
        for (unsigned m = 0; m < M; ++m) {
            double* v = &V[m * len];
            for (unsigned q = 0; q < m; ++q) {
                const double* u = &V[q * len];
                double d = 0;
                for (size_t c = 0; c < len; ++c) d += u[c] * v[c];
                for (size_t c = 0; c < len; ++c) v[c] -= d * u[c];
            }
            double nn = 0;
            for (size_t c = 0; c < len; ++c) nn += v[c] * v[c];
            nn = 1.0 / std::sqrt(nn);
            for (size_t c = 0; c < len; ++c) v[c] *= nn;
        }
    // Component c of bead i is V[m][c*n + i] -> reference layout [c][i*M + m].
    for (unsigned m = 0; m < M; ++m)
        for (unsigned i = 0; i < n; ++i)
            for (int c = 0; c < 3; ++c)
                modes[(size_t)c * n * M + (size_t)i * M + m] = (float)V[m * len + (size_t)c * n + i];
    for (unsigned m = 0; m < M; ++m) {
        const float h = (float)(h_lo + (h_hi - h_lo) * unif(seed, entity + 1000, m));
        force_const[m] = h * h;
    }
}
