// gft_oracle.cpp -- CPU restatement of the reference evolver.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP evolver in
// gftevolver_amd/csrc: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and never as the thing measured or shipped.
//
// It restates, with the same operation order and the same float/double
// conversion points, the reference's two-time-level leapfrog of
//   * Core/Source/Lattice.cpp:755-876   Lattice::evolve (slot rotation, site order)
//   * Core/Source/Model.cpp:339-375     Model::evolve (scalar leapfrog, gauge RHS)
//   * Core/Source/Model.cpp:323-337     Model::calcConstraintViolation
//   * Gradients/Source/StandardModel.cpp:72-363   transform, Laplacian, currents,
//                                                 gradient/kinetic energy, Gauss charge
//   * Gradients/Source/TwoHDM.cpp:72-116, 270-489 the same for two doublets (gaugeNum 4)
//   * WilsonLoops/Source/StandardModel.cpp:160-514 SU2Product, magnetic/electric
//                                                 forces and energies, link update, Gauss
//   * WilsonLoops/Source/StandardModel.cpp:78-157  quaternion and generator link forms
//   * WilsonLoops/Source/TwoHDM.cpp:407-490      identical to the SM for gaugeNum 4
//   * Potentials/Source/SO_N.cpp:57-87, Potentials/Source/Double_SO_N.cpp:157-368
//   * Analysers/Source/Energy.cpp:117-220, GaugeCondition.cpp:117-226 (float accumulators)
//   * BoundaryConditions/Source/Periodic.cpp:18-170 (periodic neighbours) and the
//     26-region shell order of Lattice.cpp:262-290 / BoundaryCondition.cpp:21-70,
//     which fixes the order of the analysers' float sums.
//   * InitialConditions/Source/RandomUniform.cpp:60-77 (std::mt19937 initial state).
//
// Storage type T is float (the reference's own storage, Lattice.hpp:32) or
// double (the FP64 reference).  Field arithmetic is double in both cases, as in
// the reference.  The "ref" diagnostics use float accumulators in the
// reference's site order (only meaningful for T = float); the "f64"
// diagnostics are FP64 per site with Neumaier-compensated sums.
//
// Build with -ffp-contract=off: the reference build (x86-64 baseline, -O3) has
// no FMA, so contraction would change rounding.
#include "../include/gft.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {


struct Neumaier {
    double s = 0.0, c = 0.0;
    void add(double x) {
        double t = s + x;
        if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x; else c += (x - t) + s;
        s = t;
    }
    double value() const { return s + c; }
};

struct Params {
    int Ns = 4, nd = 1;
    long nx = 0, ny = 0, nz = 0, N = 0;
    double dx = 0, dy = 0, dz = 0, dt = 0;
    double lap[3][3];     // Gradients::StandardModel::configure stencils (StandardModel.cpp:14-24)
    double est[3][3];     // energy stencils (StandardModel.cpp:35-42)
    double inv_sp[3];     // WilsonLoops inverse_sqr_spacings (StandardModel.cpp:51-53)
    double inv_sqr_dt;
    double g_sqr, gp_sqr, inv_g_sqr, inv_gp_sqr;
    float gsq_f[7];       // getSqrCouplings returns float (StandardModel.cpp:224-230)
    const gft_desc* d;
    float damping = 0.f;
    unsigned nt_damped = 0;
    int fixed = 0;        // fixed faces (bit 2a low, 2a+1 high face of axis a)
    bool gen = false;     // generator representation of the SU(2) links
    int vd = 5, nv = 15;  // reals per link direction and per site: [y, c0..c3] or [y, w1..w3]
    double tol = 1e-9;    // division-by-zero tolerance of the unit-vector conversions
    // Gauge group (WilsonLoops/Source/TwoHDM.cpp:215-242): G = 4 for the SM
    // group, 5 adds a Higgs-family U(1) (phase y' at link[off]), 7 a
    // Higgs-family SU(2) (link[off..] in the same form as the isospin SU(2)).
    int G = 4, off = 0;
    double inv_gpp_sqr = 0.0;
    // Global O(N) model (Gradients/Source/Global.cpp with the null Wilson loop):
    // N = Ns scalar components, no links, no Gauss law; so_n selects SO_N(N)
    // over Double_SO_N.
    bool global = false, so_n = true;
};

Params make_params(const gft_desc* d) {
    Params p;
    p.d = d;
    p.global = d->model == GFT_MODEL_GLOBAL;
    p.Ns = p.global ? d->num_scalar : d->model == GFT_MODEL_2HDM ? 8 : 4;
    p.nd = p.Ns / 4;
    p.so_n = p.global ? d->potential != 2 : d->model != GFT_MODEL_2HDM;
    p.nx = d->nx; p.ny = d->ny; p.nz = d->nz; p.N = p.nx * p.ny * p.nz;
    p.dx = d->dx; p.dy = d->dy; p.dz = d->dz; p.dt = d->dt;
    const double sp[3] = {d->dx, d->dy, d->dz};
    for (int a = 0; a < 3; a++) {
        p.lap[a][0] = 1.0 / (sp[a] * sp[a]);
        p.lap[a][1] = -2.0 / (sp[a] * sp[a]);
        p.lap[a][2] = 1.0 / (sp[a] * sp[a]);
        p.est[a][0] = 0.0;
        p.est[a][1] = -1.0 / sp[a];
        p.est[a][2] = 1.0 / sp[a];
        p.inv_sp[a] = 1.f / (sp[a] * sp[a]);
    }
    p.inv_sqr_dt = 1.f / (d->dt * d->dt);
    p.g_sqr = d->g * d->g;
    p.gp_sqr = d->gp * d->gp;
    p.inv_g_sqr = p.g_sqr > d->div_tol ? 1.0 / p.g_sqr : 0.0;
    p.inv_gp_sqr = p.gp_sqr > d->div_tol ? 1.0 / p.gp_sqr : 0.0;
    // getSqrCouplings returns float: g'^2 for generator 0, g^2 for 1..3, g''^2
    // for the Higgs-family generators (TwoHDM.cpp:258-267).
    p.gsq_f[0] = static_cast<float>(p.gp_sqr);
    p.gsq_f[1] = p.gsq_f[2] = p.gsq_f[3] = static_cast<float>(p.g_sqr);
    const double gpp_sqr = d->gpp * d->gpp;
    p.gsq_f[4] = p.gsq_f[5] = p.gsq_f[6] = static_cast<float>(gpp_sqr);
    p.inv_gpp_sqr = gpp_sqr > d->div_tol ? 1.0 / gpp_sqr : 0.0;
    p.damping = d->damping;
    p.nt_damped = d->nt_damped;
    p.fixed = d->fixed_faces;
    p.gen = d->gen_rep != 0;
    p.G = (d->gauge_num == 5 || d->gauge_num == 7) ? d->gauge_num : 4;
    p.off = p.gen ? 4 : 5;
    p.vd = p.off + (p.G == 5 ? 1 : p.G == 7 ? (p.gen ? 3 : 4) : 0);
    p.nv = 3 * p.vd;
    if (p.global) { p.G = 0; p.vd = p.nv = 0; }
    p.tol = d->div_tol;
    return p;
}

// A site in the one-site shell of a fixed face: Fixed::evolve (Fixed.cpp:159-240)
// leaves it alone and Fixed::postEvolveAnalysis (Fixed.cpp:245-251) and the
// skipped location analysis leave it out of the diagnostics; regions touching
// any fixed face are fixed (Lattice.cpp:293-312).
inline bool frozen(const Params& p, long s) {
    if (!p.fixed) return false;
    const long z = s % p.nz, y = (s / p.nz) % p.ny, x = s / (p.nz * p.ny);
    const long c[3] = {x, y, z}, n[3] = {p.nx, p.ny, p.nz};
    for (int a = 0; a < 3; a++) {
        if ((p.fixed >> (2 * a) & 1) && c[a] == 0) return true;
        if ((p.fixed >> (2 * a + 1) & 1) && c[a] == n[a] - 1) return true;
    }
    return false;
}

// Periodic neighbour of site s along axis a by delta in {-1, +1}
// (Periodic.cpp:50, 160-165).
inline long shift(const Params& p, long s, int a, int delta) {
    long z = s % p.nz, y = (s / p.nz) % p.ny, x = s / (p.nz * p.ny);
    if (a == 0) x = (x + delta + p.nx) % p.nx;
    else if (a == 1) y = (y + delta + p.ny) % p.ny;
    else z = (z + delta + p.nz) % p.nz;
    return (x * p.ny + y) * p.nz + z;
}

// getSU2Representation (WilsonLoops/Source/StandardModel.cpp:78-111,
// Gradients/Source/StandardModel.cpp:103-136).  Quaternion links store
// [y, c0, c1, c2, c3]; generator links store [y, w1, w2, w3] with
// U = cos|w| + i sin|w| (w/|w|).sigma, and below the division-by-zero tolerance
// the reference picks c1 = sin|w| without the conjugation sign.
template <class T>
inline void rep(const Params& p, const T* link, bool conj, double* U) {
    const int cf = conj ? -1 : 1;
    if (p.gen) {
        double wm = 0.0;
        for (int k = 1; k < 4; k++) wm += static_cast<double>(link[k]) * static_cast<double>(link[k]);
        wm = std::sqrt(wm);
        U[0] = std::cos(wm);
        U[1] = U[2] = U[3] = 0.0;
        if (wm > p.tol) {
            for (int k = 1; k < 4; k++) U[k] = cf * std::sin(wm) * static_cast<double>(link[k]) / wm;
        } else {
            U[1] = std::sin(wm);
        }
        return;
    }
    U[0] = static_cast<double>(link[1]);
    for (int k = 1; k < 4; k++) U[k] = static_cast<double>(cf * link[1 + k]);
}

// invertSU2Representation (WilsonLoops/Source/StandardModel.cpp:114-157): the
// generator components of a unit quaternion, |w| = asin|c| (pi/2 when |c|
// rounds above 1), reflected to pi - |w| when c0 < 0.  The reference returns
// them as float; storage T = double keeps them in double.
template <class T>
inline void invert_rep(const Params& p, const double* c, T* w) {
    double cm = 0.0;
    for (int k = 1; k < 4; k++) cm += c[k] * c[k];
    cm = std::sqrt(cm);
    double ph = cm > 1 ? M_PI_2 : std::asin(cm);
    if (c[0] < 0) ph = M_PI - ph;
    if (cm > p.tol) {
        for (int k = 0; k < 3; k++) w[k] = static_cast<T>(ph * c[k + 1] / cm);
    } else {
        w[0] = static_cast<T>(ph);
        w[1] = w[2] = 0;
    }
}

// SU2Product (WilsonLoops/Source/StandardModel.cpp:160-191).
inline void qmul(const double* a, const double* b, bool trace, double* out) {
    double r[4];
    r[1] = a[0] * b[1] + a[1] * b[0] - a[2] * b[3] + a[3] * b[2];
    r[2] = a[0] * b[2] + a[2] * b[0] - a[3] * b[1] + a[1] * b[3];
    r[3] = a[0] * b[3] + a[3] * b[0] - a[1] * b[2] + a[2] * b[1];
    r[0] = 0.0;
    if (trace) {
        double dot = 0.0;
        for (int k = 1; k < 4; k++) dot += a[k] * b[k];
        r[0] = a[0] * b[0] - dot;
        double nrm = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
        for (int k = 0; k < 4; k++) r[k] = r[k] / nrm;
    }
    for (int k = 0; k < 4; k++) out[k] = r[k];
}

// transform: U(1) phase on the pairs (0,1),(2,3),... then the SU(2) real
// 4x4 action per doublet (Gradients/Source/StandardModel.cpp:72-100;
// TwoHDM.cpp:72-104 for two doublets), then for G = 5 the Higgs-family phase,
// opposite on the two doublets, or for G = 7 the Higgs-family SU(2) acting on
// (f_i, f_i+1, f_i+4, f_i+5), i = 0, 2 (TwoHDM.cpp:106-141).
template <class T>
inline void transform_sm(const Params& p, const double* sf, int nd, const T* link, bool conj, double* out);

template <class T>
inline void transform(const Params& p, const double* sf, int nd, const T* link, bool conj, double* out) {
    if (p.G == 4) {
        transform_sm(p, sf, nd, link, conj, out);
        return;
    }
    double sm[8];
    transform_sm(p, sf, nd, link, conj, sm);
    const int cf = conj ? -1 : 1;
    if (p.G == 5) {
        const double y2 = static_cast<double>(link[p.off]);
        for (int c = 0; c < 8; c++) {
            const int pm = -1 + 2 * (c % 2);
            const int pm4 = 1 - 2 * (c / 4);
            out[c] = std::cos(y2) * sm[c] + pm * std::sin(pm4 * cf * y2) * sm[c - pm];
        }
        return;
    }
    double U[4];
    rep(p, link + p.off - 1, conj, U);
    for (int i = 0; i < 4; i += 2) {
        out[i] = U[0] * sm[i] - U[3] * sm[i + 1] + U[2] * sm[i + 4] - U[1] * sm[i + 5];
        out[i + 1] = U[3] * sm[i] + U[0] * sm[i + 1] + U[1] * sm[i + 4] + U[2] * sm[i + 5];
        out[i + 4] = -U[2] * sm[i] - U[1] * sm[i + 1] + U[0] * sm[i + 4] + U[3] * sm[i + 5];
        out[i + 5] = U[1] * sm[i] - U[2] * sm[i + 1] - U[3] * sm[i + 4] + U[0] * sm[i + 5];
    }
}

template <class T>
inline void transform_sm(const Params& p, const double* sf, int nd, const T* link, bool conj, double* out) {
    const int cf = conj ? -1 : 1;
    const double y = static_cast<double>(link[0]);
    const double cy = std::cos(y);
    const double sy = std::sin(cf * y);
    double h[8];
    for (int c = 0; c < 4 * nd; c++) {
        const int pm = -1 + 2 * (c % 2);
        h[c] = cy * sf[c] + pm * sy * sf[c - pm];
    }
    double U[4];
    rep(p, link, conj, U);
    for (int k = 0; k < nd; k++) {
        const double* hk = h + 4 * k;
        double* o = out + 4 * k;
        o[0] = U[0] * hk[0] - U[3] * hk[1] + U[2] * hk[2] - U[1] * hk[3];
        o[1] = U[3] * hk[0] + U[0] * hk[1] + U[1] * hk[2] + U[2] * hk[3];
        o[2] = -U[2] * hk[0] - U[1] * hk[1] + U[0] * hk[2] + U[3] * hk[3];
        o[3] = U[1] * hk[0] - U[2] * hk[1] - U[3] * hk[2] + U[0] * hk[3];
    }
}

// The four bilinears of currents and Gauss charges (StandardModel.cpp:346-358).
inline void bilinears(const double* t, const double* n, double* j) {
    j[0] = 0.25 * (t[0] * n[1] - t[1] * n[0] + t[2] * n[3] - t[3] * n[2]);
    j[1] = 0.25 * (t[0] * n[3] - t[1] * n[2] + t[2] * n[1] - t[3] * n[0]);
    j[2] = 0.25 * (-t[0] * n[2] - t[1] * n[3] + t[2] * n[0] + t[3] * n[1]);
    j[3] = 0.25 * (t[0] * n[1] - t[1] * n[0] - t[2] * n[3] + t[3] * n[2]);
}

// Higgs-family bilinears of one upper/lower pair i in {0, 2} of both doublets
// (TwoHDM.cpp:401-424 currents, 299-319 Gauss charges), accumulated into j[4..G-1].
// R is the arithmetic type of the products (double for the currents, float for
// the reference's Gauss charges, which multiply the stored floats) and A that
// of the accumulator (the reference's contribution vectors are double for the
// currents, float for the Gauss charges).
template <class R, class A>
inline void family_bilinears_add(int G, const R* t, const R* n, int i, A* j) {
    if (G == 5) {
        j[4] = static_cast<A>(j[4] + 0.25 * (t[i] * n[i + 1] - t[i + 1] * n[i] - t[i + 4] * n[i + 5] + t[i + 5] * n[i + 4]));
    } else if (G == 7) {
        j[4] = static_cast<A>(j[4] + 0.25 * (t[i] * n[i + 5] - t[i + 1] * n[i + 4] + t[i + 4] * n[i + 1] - t[i + 5] * n[i]));
        j[5] = static_cast<A>(j[5] + 0.25 * (-t[i] * n[i + 4] - t[i + 1] * n[i + 5] + t[i + 4] * n[i] + t[i + 5] * n[i + 1]));
        j[6] = static_cast<A>(j[6] + 0.25 * (t[i] * n[i + 1] - t[i + 1] * n[i] - t[i + 4] * n[i + 5] + t[i + 5] * n[i + 4]));
    }
}

// Potential derivative: SO_N (SO_N.cpp:57-75) or Double_SO_N (Double_SO_N.cpp:157-254).
void potential_derivative(const Params& p, const double* f, double* pc) {
    const gft_desc* d = p.d;
    if (p.so_n) {
        double fsq = 0.0;
        for (int c = 0; c < p.Ns; c++) fsq += f[c] * f[c];
        for (int c = 0; c < p.Ns; c++) pc[c] = d->lambda * (fsq - d->eta_sqr) * f[c];
        return;
    }
    const int n0 = 4, n1 = 4;
    double fsq[2] = {0.0, 0.0}, dot = 0.0, anti = 0.0;
    for (int fi = 0; fi < 2; fi++) {
        for (int c = 0; c < (fi ? n1 : n0); c++) fsq[fi] += f[fi * n0 + c] * f[fi * n0 + c];
        if (d->break_to_son) {
            dot = 0.0;
            for (int c = 0; c < n0; c++) dot += f[c] * f[n0 + c];
            if (d->break_to_uhalfn) {
                anti = 0.0;
                for (int c = 0; c < n0 / 2; c++)
                    anti += f[2 * c] * f[n0 + 2 * c + 1] - f[2 * c + 1] * f[n0 + 2 * c];
            }
        }
    }
    for (int c = 0; c < 8; c++) pc[c] = 0.0;
    for (int c = 0; c < n0; c++) {
        pc[c] += (-2.0 * d->m1 + 4.0 * d->l1 * fsq[0] + 2.0 * d->l3 * fsq[1]) * f[c];
        if (d->break_to_son) {
            const int f2 = n0 + c;
            pc[c] += (-d->m12r + 2.0 * d->l4p5 * dot + d->l6r * fsq[0] + d->l7r * fsq[1]) * f[f2]
                   + 2.0 * d->l6r * dot * f[c];
            if (d->break_to_uhalfn) {
                const int alt = 1 - 2 * (c % 2);
                pc[c] += (d->m12i + 2.0 * d->l4m5 * anti - 2.0 * d->l5i * dot - d->l6i * fsq[0] - d->l7i * fsq[1])
                             * alt * f[f2 + alt]
                       - 2.0 * d->l5i * anti * f[f2] - 2.0 * d->l6i * anti * f[c];
            }
        }
    }
    for (int c = 0; c < n1; c++) {
        const int f2 = n0 + c;
        pc[f2] += (-2.0 * d->m2 + 4.0 * d->l2 * fsq[1] + 2.0 * d->l3 * fsq[0]) * f[f2];
        if (d->break_to_son) {
            pc[f2] += (-d->m12r + 2.0 * d->l4p5 * dot + d->l6r * fsq[0] + d->l7r * fsq[1]) * f[c]
                    + 2.0 * d->l7r * dot * f[f2];
            if (d->break_to_uhalfn) {
                const int alt = 1 - 2 * (c % 2);
                pc[f2] += -(d->m12i + 2.0 * d->l4m5 * anti - 2.0 * d->l5i * dot - d->l6i * fsq[0] - d->l7i * fsq[1])
                              * alt * f[c + alt]
                        - 2.0 * d->l5i * anti * f[c] - 2.0 * d->l7i * anti * f[f2];
            }
        }
    }
}

// Potential energy density, reference arithmetic (SO_N.cpp:78-87; Double_SO_N.cpp:274-368).
template <class A>
A potential_energy(const Params& p, const double* f);

template <>
float potential_energy<float>(const Params& p, const double* f) {
    const gft_desc* d = p.d;
    if (p.so_n) {
        double fsq = 0.0;
        for (int c = 0; c < p.Ns; c++) fsq += f[c] * f[c];
        const float diff = static_cast<float>(fsq - d->eta_sqr);
        return static_cast<float>(0.25f * d->lambda * (diff * diff));
    }
    double fsq[2] = {0.0, 0.0};
    for (int fi = 0; fi < 2; fi++)
        for (int c = 0; c < 4; c++) fsq[fi] += f[fi * 4 + c] * f[fi * 4 + c];
    double dot = 0.0, anti = 0.0;
    if (d->break_to_son) {
        for (int c = 0; c < 4; c++) dot += f[c] * f[4 + c];
        if (d->break_to_uhalfn)
            for (int c = 0; c < 2; c++) anti += f[2 * c] * f[4 + 2 * c + 1] - f[2 * c + 1] * f[4 + 2 * c];
    }
    float v = static_cast<float>(-d->m1 * fsq[0] - d->m2 * fsq[1] + d->l1 * (fsq[0] * fsq[0])
                                 + d->l2 * (fsq[1] * fsq[1]) + d->l3 * fsq[0] * fsq[1]);
    if (d->break_to_son) {
        v = static_cast<float>(v + (-d->m12r * dot + d->l4p5 * (dot * dot) + d->l6r * fsq[0] * dot
                                    + d->l7r * fsq[1] * dot));
        if (d->break_to_uhalfn)
            v = static_cast<float>(v + (d->m12i * anti + d->l4m5 * (anti * anti) - 2.0 * d->l5i * dot * anti
                                        - d->l6i * fsq[0] * anti - d->l7i * fsq[1] * anti));
    }
    if (d->pot_norm) v += static_cast<float>(d->pot_offset);
    return v;
}

template <>
double potential_energy<double>(const Params& p, const double* f) {
    const gft_desc* d = p.d;
    if (p.so_n) {
        double fsq = 0.0;
        for (int c = 0; c < p.Ns; c++) fsq += f[c] * f[c];
        const double diff = fsq - d->eta_sqr;
        return 0.25 * d->lambda * diff * diff;
    }
    double fsq[2] = {0.0, 0.0};
    for (int fi = 0; fi < 2; fi++)
        for (int c = 0; c < 4; c++) fsq[fi] += f[fi * 4 + c] * f[fi * 4 + c];
    double dot = 0.0, anti = 0.0;
    if (d->break_to_son) {
        for (int c = 0; c < 4; c++) dot += f[c] * f[4 + c];
        if (d->break_to_uhalfn)
            for (int c = 0; c < 2; c++) anti += f[2 * c] * f[4 + 2 * c + 1] - f[2 * c + 1] * f[4 + 2 * c];
    }
    double v = -d->m1 * fsq[0] - d->m2 * fsq[1] + d->l1 * fsq[0] * fsq[0] + d->l2 * fsq[1] * fsq[1]
             + d->l3 * fsq[0] * fsq[1];
    if (d->break_to_son) {
        v += -d->m12r * dot + d->l4p5 * dot * dot + d->l6r * fsq[0] * dot + d->l7r * fsq[1] * dot;
        if (d->break_to_uhalfn)
            v += d->m12i * anti + d->l4m5 * anti * anti - 2.0 * d->l5i * dot * anti - d->l6i * fsq[0] * anti
               - d->l7i * fsq[1] * anti;
    }
    if (d->pot_norm) v += d->pot_offset;
    return v;
}

// Per-site energy densities in the order of Energy::preEvolveLocationAnalysis
// (Energy.cpp:126-132): potential, gradient, kinetic, magnetic, electric.
template <class A>
struct Dens { A v[5]; };

// One site-update of the global O(N) model: the plain Laplacian
// (Gradients/Source/Global.cpp:118-135, component outer, axis, then stencil),
// the potential, Model::evolve's scalar leapfrog (Model.cpp:345-356); energy
// densities of Global.cpp:84-115 (float derivative and powf sums) with zero
// magnetic and electric parts (the null Wilson loop, WilsonLoop.cpp:57-65).
template <class T, class A>
void site_update_global(const Params& p, T* S, long s, unsigned now, bool store_energy, double damping_step,
                        Dens<A>* dens) {
    const int Ns = p.Ns;
    const long N = p.N;
    const unsigned past = !now;
    const T* Sn = S + static_cast<long>(now) * N * Ns;
    T* sp_ = S + static_cast<long>(past) * N * Ns + s * Ns;
    const T* sn = Sn + s * Ns;
    long nb[3][3];
    for (int a = 0; a < 3; a++) { nb[a][0] = shift(p, s, a, -1); nb[a][1] = s; nb[a][2] = shift(p, s, a, +1); }
    double phi[8];
    for (int c = 0; c < Ns; c++) phi[c] = static_cast<double>(sn[c]);
    double pot[8];
    potential_derivative(p, phi, pot);
    double der[8];
    for (int c = 0; c < Ns; c++) {
        der[c] = 0.0;
        for (int a = 0; a < 3; a++)
            for (int st = 0; st < 3; st++) der[c] += p.lap[a][st] * static_cast<double>(Sn[nb[a][st] * Ns + c]);
    }
    if (store_energy && dens) {
        const T* spast = S + static_cast<long>(past) * N * Ns + s * Ns;
        if constexpr (sizeof(A) == sizeof(float)) {
            float ge = 0.f;
            for (int c = 0; c < Ns; c++)
                for (int a = 0; a < 3; a++) {
                    float dv = 0.f;
                    for (int st = 0; st < 3; st++)
                        dv = static_cast<float>(dv + p.est[a][st] * static_cast<float>(Sn[nb[a][st] * Ns + c]));
                    ge += ::powf(dv, 2);
                }
            float ke = 0.f;
            for (int c = 0; c < Ns; c++)
                ke += ::powf(static_cast<float>((static_cast<float>(sn[c]) - static_cast<float>(spast[c])) / p.dt), 2);
            dens->v[0] = potential_energy<float>(p, phi);
            dens->v[1] = 0.5f * ge;
            dens->v[2] = 0.5f * ke;
            dens->v[3] = dens->v[4] = 0.f;
        } else {
            double ge = 0.0, ke = 0.0;
            for (int c = 0; c < Ns; c++)
                for (int a = 0; a < 3; a++) {
                    const double dv = p.est[a][1] * phi[c] + p.est[a][2] * static_cast<double>(Sn[nb[a][2] * Ns + c]);
                    ge += dv * dv;
                }
            for (int c = 0; c < Ns; c++) {
                const double dd = (static_cast<double>(sn[c]) - static_cast<double>(spast[c])) / p.dt;
                ke += dd * dd;
            }
            dens->v[0] = potential_energy<double>(p, phi);
            dens->v[1] = 0.5 * ge;
            dens->v[2] = 0.5 * ke;
            dens->v[3] = dens->v[4] = 0.0;
        }
    }
    for (int c = 0; c < Ns; c++) {
        const double rhs = p.dt * p.dt * (der[c] - pot[c])
                         - damping_step * (static_cast<double>(sn[c]) - static_cast<double>(sp_[c]));
        sp_[c] = static_cast<T>(2 * static_cast<double>(sn[c]) - static_cast<double>(sp_[c]) + rhs);
    }
}

// One site-update of the two-slot leapfrog.  Reads slot `now` everywhere and
// slot `past` at s; writes slot `past` at s.
template <class T, class A>
void site_update(const Params& p, T* S, T* V, long s, unsigned now, bool store_energy, double damping_step,
                 Dens<A>* dens, unsigned long* panics) {
    const int Ns = p.Ns, nd = p.nd;
    const long N = p.N;
    const unsigned past = !now;
    const T* Sn = S + static_cast<long>(now) * N * Ns;
    const T* Vn = V + static_cast<long>(now) * N * p.nv;
    T* sp_ = S + static_cast<long>(past) * N * Ns + s * Ns;
    T* vp_ = V + static_cast<long>(past) * N * p.nv + s * p.nv;
    const T* sn = Sn + s * Ns;
    const T* vn = Vn + s * p.nv;

    long xm[3], xp[3], xpm[3][3];
    for (int a = 0; a < 3; a++) { xm[a] = shift(p, s, a, -1); xp[a] = shift(p, s, a, +1); }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) xpm[i][j] = (i == j) ? xp[i] : shift(p, xp[i], j, -1);

    double phi[8];
    for (int c = 0; c < Ns; c++) phi[c] = static_cast<double>(sn[c]);

    // Potential (Model::calcPotentialContributions).
    double pot[8];
    potential_derivative(p, phi, pot);

    // Covariant Laplacian (Gradients calcDerivatives, stencil index outer, axis inner).
    double der[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int st = 0; st < 3; st++) {
        for (int a = 0; a < 3; a++) {
            double tf[8];
            if (st == 0) {
                double f[8];
                const T* q = Sn + xm[a] * Ns;
                for (int c = 0; c < Ns; c++) f[c] = static_cast<double>(q[c]);
                transform(p, f, nd, Vn + xm[a] * p.nv + a * p.vd, false, tf);
            } else if (st == 1) {
                for (int c = 0; c < Ns; c++) tf[c] = phi[c];
            } else {
                double f[8];
                const T* q = Sn + xp[a] * Ns;
                for (int c = 0; c < Ns; c++) f[c] = static_cast<double>(q[c]);
                transform(p, f, nd, vn + a * p.vd, true, tf);
            }
            for (int c = 0; c < Ns; c++) der[c] += p.lap[a][st] * tf[c];
        }
    }

    // Currents (Gradients calcCurrents).
    const int G = p.G;
    double cur[21];
    for (int a = 0; a < 3; a++) {
        double tf[8], nb[8];
        transform(p, phi, nd, vn + a * p.vd, false, tf);
        const T* q = Sn + xp[a] * Ns;
        for (int c = 0; c < Ns; c++) nb[c] = static_cast<double>(q[c]);
        if (nd == 1) {
            bilinears(tf, nb, cur + G * a);
        } else {
            double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < 2; k++) {
                double j[4];
                bilinears(tf + 4 * k, nb + 4 * k, j);
                for (int m = 0; m < 4; m++) acc[m] += j[m];
                family_bilinears_add(G, tf, nb, 2 * k, acc);
            }
            for (int m = 0; m < G; m++) cur[G * a + m] = acc[m];
        }
    }

    // Magnetic (WilsonLoops calcMagneticContributions).
    double mag[21];
    for (int k = 0; k < 21; k++) mag[k] = 0.0;
    float stored_mag = 0.f;   // mutable float member, StandardModelLoops.hpp:28
    double mag_e = 0.0;
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            if (j == i) continue;
            const T* L_x = vn;
            const T* L_xpi = Vn + xp[i] * p.nv;
            const T* L_xpj = Vn + xp[j] * p.nv;
            const T* L_xmj = Vn + xm[j] * p.nv;
            const T* L_xpimj = Vn + xpm[i][j] * p.nv;
            const double th_loc = -static_cast<double>(L_xpi[j * p.vd]) + static_cast<double>(L_xpj[i * p.vd])
                                + static_cast<double>(L_x[j * p.vd]) - static_cast<double>(L_x[i * p.vd]);
            const double th_nb = -static_cast<double>(L_xpimj[j * p.vd]) - static_cast<double>(L_xmj[i * p.vd])
                               + static_cast<double>(L_xmj[j * p.vd]) + static_cast<double>(L_x[i * p.vd]);
            mag[G * i] += (std::sin(th_loc) - std::sin(th_nb)) * p.inv_sp[j];

            double P1[4], P2[4], Um[4];
            rep(p, L_x + i * p.vd, true, P1);
            rep(p, L_x + j * p.vd, false, Um);
            qmul(Um, P1, true, P1);
            rep(p, L_xpj + i * p.vd, false, Um);
            qmul(Um, P1, true, P1);
            rep(p, L_xpi + j * p.vd, true, Um);
            qmul(Um, P1, store_energy, P1);

            rep(p, L_xpimj + j * p.vd, true, P2);
            rep(p, L_xmj + i * p.vd, true, Um);
            qmul(Um, P2, true, P2);
            rep(p, L_xmj + j * p.vd, false, Um);
            qmul(Um, P2, true, P2);
            rep(p, L_x + i * p.vd, false, Um);
            qmul(Um, P2, false, P2);

            for (int k = 1; k < 4; k++) mag[G * i + k] += (P1[k] - P2[k]) * p.inv_sp[j];

            if (store_energy) {
                const double e = 2.0 * p.inv_sp[i] * p.inv_sp[j]
                               * (p.inv_g_sqr * (1.0 - P1[0]) + p.inv_gp_sqr * (1.0 - std::cos(th_loc)));
                stored_mag = static_cast<float>(stored_mag + e);
                mag_e += e;
            }

            // Higgs-family loops (TwoHDM.cpp:496-552), after the SM ones.
            const int o = p.off;
            if (G == 5) {
                const double f_loc = -static_cast<double>(L_xpi[j * p.vd + o]) + static_cast<double>(L_xpj[i * p.vd + o])
                                   + static_cast<double>(L_x[j * p.vd + o]) - static_cast<double>(L_x[i * p.vd + o]);
                const double f_nb = -static_cast<double>(L_xpimj[j * p.vd + o]) - static_cast<double>(L_xmj[i * p.vd + o])
                                  + static_cast<double>(L_xmj[j * p.vd + o]) + static_cast<double>(L_x[i * p.vd + o]);
                mag[G * i + 4] += (std::sin(f_loc) - std::sin(f_nb)) * p.inv_sp[j];
                if (store_energy) {
                    const double e = 2.0 * p.inv_sp[i] * p.inv_sp[j] * p.inv_gpp_sqr * (1.0 - std::cos(f_loc));
                    stored_mag = static_cast<float>(stored_mag + e);
                    mag_e += e;
                }
            } else if (G == 7) {
                rep(p, L_x + i * p.vd + o - 1, true, P1);
                rep(p, L_x + j * p.vd + o - 1, false, Um);
                qmul(Um, P1, true, P1);
                rep(p, L_xpj + i * p.vd + o - 1, false, Um);
                qmul(Um, P1, true, P1);
                rep(p, L_xpi + j * p.vd + o - 1, true, Um);
                qmul(Um, P1, store_energy, P1);
                rep(p, L_xpimj + j * p.vd + o - 1, true, P2);
                rep(p, L_xmj + i * p.vd + o - 1, true, Um);
                qmul(Um, P2, true, P2);
                rep(p, L_xmj + j * p.vd + o - 1, false, Um);
                qmul(Um, P2, true, P2);
                rep(p, L_x + i * p.vd + o - 1, false, Um);
                qmul(Um, P2, false, P2);
                for (int k = 1; k < 4; k++) mag[G * i + 3 + k] += (P1[k] - P2[k]) * p.inv_sp[j];
                if (store_energy) {
                    const double e = 2.0 * p.inv_sp[i] * p.inv_sp[j] * p.inv_gpp_sqr * (1.0 - P1[0]);
                    stored_mag = static_cast<float>(stored_mag + e);
                    mag_e += e;
                }
            }
        }
    }

    // Electric (WilsonLoops calcElectricContributions).
    double el[21];
    float stored_el = 0.f;
    double el_e = 0.0;
    const T* vpast = V + static_cast<long>(past) * N * p.nv + s * p.nv;
    for (int a = 0; a < 3; a++) {
        const double la = static_cast<double>(vn[a * p.vd]) - static_cast<double>(vpast[a * p.vd]);
        el[G * a] = std::sin(la);
        double Up[4], Um[4];
        rep(p, vpast + a * p.vd, true, Up);
        rep(p, vn + a * p.vd, false, Um);
        qmul(Um, Up, store_energy, Up);
        for (int k = 1; k < 4; k++) el[G * a + k] = Up[k];
        if (store_energy) {
            const double e = 4.0 * p.inv_sp[a] * (p.inv_g_sqr * (1.0 - Up[0]) + p.inv_gp_sqr * (1.0 - std::cos(la)));
            stored_el = static_cast<float>(stored_el + e);
            el_e += e;
        }
        // Higgs family (TwoHDM.cpp:600-634).
        const int o = p.off;
        if (G == 5) {
            const double lf = static_cast<double>(vn[a * p.vd + o]) - static_cast<double>(vpast[a * p.vd + o]);
            el[G * a + 4] = std::sin(lf);
            if (store_energy) {
                const double e = 4.0 * p.inv_sp[a] * p.inv_gpp_sqr * (1.0 - std::cos(lf));
                stored_el = static_cast<float>(stored_el + e);
                el_e += e;
            }
        } else if (G == 7) {
            rep(p, vpast + a * p.vd + o - 1, true, Up);
            rep(p, vn + a * p.vd + o - 1, false, Um);
            qmul(Um, Up, store_energy, Up);
            for (int k = 1; k < 4; k++) el[G * a + 3 + k] = Up[k];
            if (store_energy) {
                const double e = 4.0 * p.inv_sp[a] * p.inv_gpp_sqr * (1.0 - Up[0]);
                stored_el = static_cast<float>(stored_el + e);
                el_e += e;
            }
        }
    }

    // Energy densities (before the update), Energy::preEvolveLocationAnalysis.
    if (store_energy && dens) {
        const double spc[3] = {p.dx, p.dy, p.dz};
        (void)spc;
        if constexpr (sizeof(A) == sizeof(float)) {
            // Reference float arithmetic.
            float pe = potential_energy<float>(p, phi);
            // Gradients calcGradientEnergy (StandardModel.cpp:192-235).
            float derv[24];
            for (int k = 0; k < 3 * Ns; k++) derv[k] = 0.f;
            for (int st = 1; st < 3; st++) {
                for (int a = 0; a < 3; a++) {
                    double tf[8];
                    const T* q = (st == 1) ? sn : Sn + xp[a] * Ns;
                    for (int c = 0; c < Ns; c++) tf[c] = static_cast<double>(q[c]);
                    if (st == 1) {
                        double tmp[8];
                        transform(p, tf, nd, vn + a * p.vd, false, tmp);
                        for (int c = 0; c < Ns; c++) tf[c] = tmp[c];
                    }
                    for (int c = 0; c < Ns; c++)
                        derv[a * Ns + c] = static_cast<float>(derv[a * Ns + c] + p.est[a][st] * tf[c]);
                }
            }
            float ge = 0.f;
            for (int k = 0; k < 3 * Ns; k++)
                ge = static_cast<float>(ge + static_cast<double>(derv[k]) * static_cast<double>(derv[k]));
            ge = 0.5f * ge;
            // calcKineticEnergy (StandardModel.cpp:238-248), float arithmetic.
            float ke = 0.f;
            const T* spast = S + static_cast<long>(past) * N * Ns + s * Ns;
            const float dtf = static_cast<float>(p.dt);
            for (int c = 0; c < Ns; c++) {
                const float dd = (static_cast<float>(sn[c]) - static_cast<float>(spast[c])) / dtf;
                ke += dd * dd;
            }
            ke = 0.5f * ke;
            const float me = stored_mag;
            const float ee = static_cast<float>(p.inv_sqr_dt * stored_el);
            dens->v[0] = pe; dens->v[1] = ge; dens->v[2] = ke; dens->v[3] = me; dens->v[4] = ee;
        } else {
            double pe = potential_energy<double>(p, phi);
            double ge = 0.0;
            for (int a = 0; a < 3; a++) {
                double tf[8], nb[8];
                transform(p, phi, nd, vn + a * p.vd, false, tf);
                const T* q = Sn + xp[a] * Ns;
                for (int c = 0; c < Ns; c++) nb[c] = static_cast<double>(q[c]);
                for (int c = 0; c < Ns; c++) {
                    const double dd = p.est[a][1] * tf[c] + p.est[a][2] * nb[c];
                    ge += dd * dd;
                }
            }
            ge *= 0.5;
            double ke = 0.0;
            const T* spast = S + static_cast<long>(past) * N * Ns + s * Ns;
            for (int c = 0; c < Ns; c++) {
                const double dd = (static_cast<double>(sn[c]) - static_cast<double>(spast[c])) / p.dt;
                ke += dd * dd;
            }
            ke *= 0.5;
            dens->v[0] = pe; dens->v[1] = ge; dens->v[2] = ke; dens->v[3] = mag_e;
            dens->v[4] = p.inv_sqr_dt * el_e;
        }
    }

    // Model::evolve.
    for (int c = 0; c < Ns; c++) {
        const double rhs = p.dt * p.dt * (der[c] - pot[c])
                         - damping_step * (static_cast<double>(sn[c]) - static_cast<double>(sp_[c]));
        sp_[c] = static_cast<T>(2 * static_cast<double>(sn[c]) - static_cast<double>(sp_[c]) + rhs);
    }
    double R[21];
    for (int c = 0; c < 3 * G; c++)
        R[c] = (1.0 - damping_step) * el[c]
             + p.dt * p.dt * (mag[c] + static_cast<double>(p.gsq_f[c % G]) * cur[c]);

    // WilsonLoops::StandardModel::evolve, TwoHDM::evolve (TwoHDM.cpp:640-755).
    for (int a = 0; a < 3; a++) {
        const double* Ra = R + G * a;
        vp_[a * p.vd] = static_cast<T>(std::asin(Ra[0]) + static_cast<double>(vn[a * p.vd]));
        if (Ra[0] * Ra[0] > 1) __atomic_add_fetch(panics, 1UL, __ATOMIC_RELAXED);
        double cms = 0.0;
        for (int k = 1; k < 4; k++) cms += Ra[k] * Ra[k];
        if (cms > 1) __atomic_add_fetch(panics, 1UL, __ATOMIC_RELAXED);
        const double u[4] = {std::sqrt(1.0 - cms), Ra[1], Ra[2], Ra[3]};
        double Up[4];
        rep(p, vn + a * p.vd, false, Up);
        qmul(u, Up, true, Up);
        if (p.gen) invert_rep(p, Up, vp_ + a * p.vd + 1);
        else for (int k = 0; k < 4; k++) vp_[a * p.vd + 1 + k] = static_cast<T>(Up[k]);
        const int o = p.off;
        if (G == 5) {
            vp_[a * p.vd + o] = static_cast<T>(std::asin(Ra[4]) + static_cast<double>(vn[a * p.vd + o]));
            if (Ra[4] * Ra[4] > 1) __atomic_add_fetch(panics, 1UL, __ATOMIC_RELAXED);
        } else if (G == 7) {
            double cf = 0.0;
            for (int k = 4; k < 7; k++) cf += Ra[k] * Ra[k];
            if (cf > 1) __atomic_add_fetch(panics, 1UL, __ATOMIC_RELAXED);
            const double uf[4] = {std::sqrt(1.0 - cf), Ra[4], Ra[5], Ra[6]};
            rep(p, vn + a * p.vd + o - 1, false, Up);
            qmul(uf, Up, true, Up);
            if (p.gen) invert_rep(p, Up, vp_ + a * p.vd + o);
            else for (int k = 0; k < 4; k++) vp_[a * p.vd + o + k] = static_cast<T>(Up[k]);
        }
    }
}

// Gauss-law violation at s after the update (Model::calcConstraintViolation):
// slot `now` holds t, slot `!now` holds t+dt.  Outputs per generator the
// violation and, for the FP64 flavour, the scale sum_i |electric terms| + g_a^2 |rho|.
template <class T, class A>
void site_gauss(const Params& p, const T* S, const T* V, long s, unsigned now, A* viol, double* scale) {
    const int Ns = p.Ns, nd = p.nd;
    const long N = p.N;
    const unsigned fut = !now;
    const T* sn = S + static_cast<long>(now) * N * Ns + s * Ns;
    const T* sf = S + static_cast<long>(fut) * N * Ns + s * Ns;
    const T* Vn = V + static_cast<long>(now) * N * p.nv;
    const T* Vf = V + static_cast<long>(fut) * N * p.nv;
    const int G = p.G;
    A grad[7] = {0, 0, 0, 0, 0, 0, 0};
    if constexpr (sizeof(A) == sizeof(float)) {
        // Gradients calcConstraintContributions: float products (StandardModel.cpp:250-270).
        for (int k = 0; k < nd; k++) {
            const int o = 4 * k;
            const float a0 = sn[o], a1 = sn[o + 1], a2 = sn[o + 2], a3 = sn[o + 3];
            const float b0 = sf[o], b1 = sf[o + 1], b2 = sf[o + 2], b3 = sf[o + 3];
            const float c0 = static_cast<float>(0.25 * (a0 * b1 - a1 * b0 + a2 * b3 - a3 * b2));
            const float c1 = static_cast<float>(0.25 * (a0 * b3 - a1 * b2 + a2 * b1 - a3 * b0));
            const float c2 = static_cast<float>(0.25 * (-a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1));
            const float c3 = static_cast<float>(0.25 * (a0 * b1 - a1 * b0 - a2 * b3 + a3 * b2));
            if (nd == 1) { grad[0] = c0; grad[1] = c1; grad[2] = c2; grad[3] = c3; }
            else {
                // TwoHDM.cpp:255-267: contribution += 0.25*(...) evaluated in double.
                const double e0 = 0.25 * (a0 * b1 - a1 * b0 + a2 * b3 - a3 * b2);
                const double e1 = 0.25 * (a0 * b3 - a1 * b2 + a2 * b1 - a3 * b0);
                const double e2 = 0.25 * (-a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1);
                const double e3 = 0.25 * (a0 * b1 - a1 * b0 - a2 * b3 + a3 * b2);
                grad[0] = static_cast<float>(grad[0] + e0); grad[1] = static_cast<float>(grad[1] + e1);
                grad[2] = static_cast<float>(grad[2] + e2); grad[3] = static_cast<float>(grad[3] + e3);
                // Higgs-family charges (TwoHDM.cpp:299-319), float products.
                float af[8], bf[8];
                for (int m = 0; m < 8; m++) { af[m] = static_cast<float>(sn[m]); bf[m] = static_cast<float>(sf[m]); }
                family_bilinears_add(G, af, bf, 2 * k, grad);
            }
        }
    } else {
        double t8[8], n8[8];
        for (int m = 0; m < 4 * nd; m++) { t8[m] = static_cast<double>(sn[m]); n8[m] = static_cast<double>(sf[m]); }
        for (int k = 0; k < nd; k++) {
            double j[4];
            bilinears(t8 + 4 * k, n8 + 4 * k, j);
            for (int m = 0; m < 4; m++) grad[m] += j[m];
            if (nd == 2) family_bilinears_add(G, t8, n8, 2 * k, grad);
        }
    }
    // WilsonLoops calcConstraintContributions (StandardModel.cpp:274-322).
    A elc[7] = {0, 0, 0, 0, 0, 0, 0};
    double divabs[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 3; a++) {
        const long xm = shift(p, s, a, -1);
        const T* Ln = Vn + s * p.nv + a * p.vd;
        const T* Lf = Vf + s * p.nv + a * p.vd;
        const T* Mn = Vn + xm * p.nv + a * p.vd;
        const T* Mf = Vf + xm * p.nv + a * p.vd;
        const double la = static_cast<double>(Lf[0]) - static_cast<double>(Ln[0]);
        const double na = static_cast<double>(Mf[0]) - static_cast<double>(Mn[0]);
        const double d0 = (std::sin(la) - std::sin(na)) * p.inv_sp[a];
        elc[0] = static_cast<A>(elc[0] + d0);
        divabs[0] += (std::fabs(std::sin(la)) + std::fabs(std::sin(na))) * p.inv_sp[a];
        double P1[4], P2[4], Um[4];
        rep(p, Lf, false, P1);
        rep(p, Ln, true, Um);
        qmul(Um, P1, false, P1);
        rep(p, Mn, true, P2);
        rep(p, Mf, false, Um);
        qmul(Um, P2, false, P2);
        for (int k = 1; k < 4; k++) {
            const double dk = (P1[k] - P2[k]) * p.inv_sp[a];
            elc[k] += static_cast<A>(dk);
            divabs[k] += (std::fabs(P1[k]) + std::fabs(P2[k])) * p.inv_sp[a];
        }
        // Higgs family (TwoHDM.cpp:352-386).
        const int o = p.off;
        if (G == 5) {
            const double lf = static_cast<double>(Lf[o]) - static_cast<double>(Ln[o]);
            const double nf = static_cast<double>(Mf[o]) - static_cast<double>(Mn[o]);
            elc[4] = static_cast<A>(elc[4] + (std::sin(lf) - std::sin(nf)) * p.inv_sp[a]);
            divabs[4] += (std::fabs(std::sin(lf)) + std::fabs(std::sin(nf))) * p.inv_sp[a];
        } else if (G == 7) {
            rep(p, Lf + o - 1, false, P1);
            rep(p, Ln + o - 1, true, Um);
            qmul(Um, P1, false, P1);
            rep(p, Mn + o - 1, true, P2);
            rep(p, Mf + o - 1, false, Um);
            qmul(Um, P2, false, P2);
            for (int k = 1; k < 4; k++) {
                const double dk = (P1[k] - P2[k]) * p.inv_sp[a];
                elc[3 + k] += static_cast<A>(dk);
                divabs[3 + k] += (std::fabs(P1[k]) + std::fabs(P2[k])) * p.inv_sp[a];
            }
        }
    }
    for (int k = 0; k < G; k++) {
        if constexpr (sizeof(A) == sizeof(float)) viol[k] = elc[k] - p.gsq_f[k] * grad[k];
        else viol[k] = elc[k] - static_cast<double>(p.gsq_f[k]) * grad[k];
        // Scale of the relative violation: the magnitudes of the terms that cancel.
        if (scale) scale[k] = divabs[k] + std::fabs(static_cast<double>(p.gsq_f[k]) * static_cast<double>(grad[k]));
    }
}

// Site order of the reference step: the 26 periodic boundary regions in the
// order built by Lattice::initBoundaryCondition (Lattice.cpp:262-290), each
// with the loop limits of BoundaryCondition::determineResponsibilities
// (BoundaryCondition.cpp:21-70), then the interior [1, n-1)^3 (Lattice.cpp:793-803).
std::vector<long> reference_site_order(long nx, long ny, long nz) {
    std::vector<long> order;
    order.reserve(nx * ny * nz);
    const long n[3] = {nx, ny, nz};
    for (unsigned wb = 0; wb < 26; wb++) {
        int bv[3] = {0, 0, 0};
        const int alt = -1 + 2 * static_cast<int>(wb % 2);
        const unsigned sc = wb / 2;
        const int salt = 1 - 2 * static_cast<int>(sc % 2);
        if (wb < 6) bv[sc] = alt;
        else if (wb < 18) { bv[1 + wb / 10] = alt; bv[wb / 14] = salt; }
        else { bv[0] = -1 + 2 * static_cast<int>(wb / 22); bv[1] = salt; bv[2] = alt; }
        long lo[3], hi[3];
        bool empty = false;
        for (int a = 0; a < 3; a++) {
            if (bv[a] == -1) { lo[a] = 0; hi[a] = 1; }
            else if (bv[a] == 0) { lo[a] = 1; hi[a] = n[a] < 2 ? 1 : n[a] - 1; }
            else { lo[a] = n[a] < 2 ? 1 : n[a] - 1; hi[a] = n[a]; }
            if (lo[a] >= hi[a]) empty = true;
        }
        if (empty) continue;
        for (long x = lo[0]; x < hi[0]; x++)
            for (long y = lo[1]; y < hi[1]; y++)
                for (long z = lo[2]; z < hi[2]; z++) order.push_back((x * ny + y) * nz + z);
    }
    for (long x = 1; x < nx - 1; x++)
        for (long y = 1; y < ny - 1; y++)
            for (long z = 1; z < nz - 1; z++) order.push_back((x * ny + y) * nz + z);
    return order;
}

// Columns of a reference row: 6 energies, GFT_MAX_EQ integrated |G|, GFT_MAX_EQ max |G|.
constexpr int kRefCols = 6 + 2 * GFT_MAX_EQ;

template <class T>
int evolve(const gft_desc* d, T* S, T* V, unsigned time_iter, unsigned nsteps, unsigned energy_freq,
           unsigned gauss_freq, double* ref_rows, gft_obs* rows, unsigned max_rows, unsigned* nrows,
           int nthreads, unsigned long* panics_out) {
    using A = T;
    if (!d || !S || (!V && d->model != GFT_MODEL_GLOBAL)) return GFT_EINVAL;
    if (d->nx < 2 || d->ny < 2 || d->nz < 2) return GFT_EINVAL;
    const Params p = make_params(d);
    const std::vector<long> order = reference_site_order(p.nx, p.ny, p.nz);
    if (static_cast<long>(order.size()) != p.N) return GFT_EINVAL;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    std::vector<Dens<A>> dens;
    std::vector<A> gviol;
    std::vector<double> gscale;
    unsigned row = 0;
    unsigned long panics = 0;
    const double spv = p.dx * p.dy * p.dz;
    for (unsigned k = 0; k < nsteps; k++) {
        const unsigned it = time_iter + k;
        const unsigned now = (it + 1) % 2;
        const bool meas_e = energy_freq && (it % energy_freq == 0);
        const bool meas_g = gauss_freq && (it % gauss_freq == 0);
        const float cur_damp = it < p.nt_damped ? p.damping : 0.f;
        const double damping_step = 0.5 * static_cast<double>(cur_damp) * p.dt;
        if (meas_e) dens.assign(p.N, Dens<A>{});
#pragma omp parallel for schedule(static)
        for (long s = 0; s < p.N; s++) {
            if (frozen(p, s)) continue;
            if (p.global)
                site_update_global<T, A>(p, S, s, now, meas_e, damping_step, meas_e ? &dens[s] : nullptr);
            else
                site_update<T, A>(p, S, V, s, now, meas_e, damping_step, meas_e ? &dens[s] : nullptr, &panics);
        }
        const int G = p.G;
        if (meas_g && G > 0) {
            gviol.assign(p.N * G, A(0));
            gscale.assign(p.N * G, 0.0);
#pragma omp parallel for schedule(static)
            for (long s = 0; s < p.N; s++)
                if (!frozen(p, s)) site_gauss<T, A>(p, S, V, s, now, &gviol[s * G], &gscale[s * G]);
        }
        if ((meas_e || meas_g) && row < max_rows) {
            gft_obs* o = rows ? &rows[row] : nullptr;
            double* rr = ref_rows ? &ref_rows[static_cast<size_t>(row) * kRefCols] : nullptr;
            if (o) { std::memset(o, 0, sizeof(*o)); o->time_iter = it; o->num_eq = G; }
            if (rr) for (int m = 0; m < kRefCols; m++) rr[m] = 0.0;
            if (meas_e) {
                // Reference accumulation: float sums in the reference site order (Energy.cpp:132,156,181).
                A acc[6] = {0, 0, 0, 0, 0, 0};
                Neumaier nm[6];
                for (long s : order) {
                    if (frozen(p, s)) continue;
                    const Dens<A>& dd = dens[s];
                    const A tot = dd.v[0] + dd.v[1] + dd.v[2] + dd.v[3] + dd.v[4];
                    acc[0] += tot;
                    for (int m = 0; m < 5; m++) acc[m + 1] += dd.v[m];
                    double dtot = 0.0;
                    for (int m = 0; m < 5; m++) { nm[m + 1].add(static_cast<double>(dd.v[m])); dtot += static_cast<double>(dd.v[m]); }
                    nm[0].add(dtot);
                }
                if (rr) for (int m = 0; m < 6; m++) rr[m] = static_cast<double>(acc[m]) * p.dx * p.dy * p.dz;
                if (o) {
                    o->has_energy = 1;
                    for (int m = 0; m < 6; m++) o->energy[m] = nm[m].value() * spv;
                }
            }
            if (meas_g) {
                A integ[7] = {0, 0, 0, 0, 0, 0, 0}, mx[7] = {0, 0, 0, 0, 0, 0, 0};
                Neumaier ni[7], ns[7];
                double dmx[7] = {0, 0, 0, 0, 0, 0, 0};
                for (long s : order) {
                    if (frozen(p, s)) continue;
                    for (int m = 0; m < G; m++) {
                        const A v = std::abs(gviol[s * G + m]);
                        integ[m] += v;
                        if (v > mx[m]) mx[m] = v;
                        ni[m].add(static_cast<double>(v));
                        ns[m].add(gscale[s * G + m]);
                        if (static_cast<double>(v) > dmx[m]) dmx[m] = static_cast<double>(v);
                    }
                }
                if (rr) {
                    for (int m = 0; m < G; m++) {
                        // GaugeCondition.cpp:175: integrated *= dx*dy*dz in the accumulator type.
                        const A scaled = static_cast<A>(integ[m] * (p.dx * p.dy * p.dz));
                        rr[6 + m] = static_cast<double>(scaled);
                        rr[6 + GFT_MAX_EQ + m] = static_cast<double>(mx[m]);
                    }
                }
                if (o) {
                    o->has_gauss = 1;
                    for (int m = 0; m < G; m++) {
                        o->gauss_int[m] = ni[m].value() * spv;
                        o->gauss_max[m] = dmx[m];
                        o->gauss_scale[m] = ns[m].value() * spv;
                    }
                }
            }
            row++;
        }
    }
    if (nrows) *nrows = row;
    if (panics_out) *panics_out = panics;
    return GFT_OK;
}

}  // namespace

extern "C" {

// Two-slot evolution with FP32 storage (the reference's own arithmetic).
// ref_rows: [rows][6 + 2 * GFT_MAX_EQ] = 6 energy columns, then GFT_MAX_EQ
// integrated |G| and GFT_MAX_EQ max |G| (the first num_eq of each used) exactly
// as the reference accumulates them; rows: FP64 diagnostics of the same states.
int oracle_evolve_f32(const gft_desc* d, float* scalar, float* vec, unsigned time_iter, unsigned nsteps,
                      unsigned energy_freq, unsigned gauss_freq, double* ref_rows, gft_obs* rows,
                      unsigned max_rows, unsigned* nrows, int nthreads, unsigned long* panics) {
    return evolve<float>(d, scalar, vec, time_iter, nsteps, energy_freq, gauss_freq, ref_rows, rows, max_rows,
                         nrows, nthreads, panics);
}

// Two-slot evolution with FP64 storage: the FP64 reference the GPU is held to.
int oracle_evolve_f64(const gft_desc* d, double* scalar, double* vec, unsigned time_iter, unsigned nsteps,
                      unsigned energy_freq, unsigned gauss_freq, double* ref_rows, gft_obs* rows,
                      unsigned max_rows, unsigned* nrows, int nthreads, unsigned long* panics) {
    return evolve<double>(d, scalar, vec, time_iter, nsteps, energy_freq, gauss_freq, ref_rows, rows, max_rows,
                          nrows, nthreads, panics);
}

// Per-site reference energy densities of the state before step time_iter
// (potential, gradient, kinetic, magnetic, electric; float arithmetic as in
// Energy::preEvolveLocationAnalysis).  Fields are not modified.
int oracle_energy_density_f32(const gft_desc* d, const float* scalar, const float* vec, unsigned time_iter,
                              float* dens) {
    const Params p = make_params(d);
    std::vector<float> S(scalar, scalar + 2 * p.N * p.Ns), V(vec, vec + 2 * p.N * p.nv);
    const unsigned now = (time_iter + 1) % 2;
    unsigned long panics = 0;
    std::vector<Dens<float>> dd(p.N);
    // Each site writes only its own past slot, so updating the copy in any
    // order leaves every other site's inputs untouched.
    for (long s = 0; s < p.N; s++) {
        if (p.global) site_update_global<float, float>(p, S.data(), s, now, true, 0.0, &dd[s]);
        else site_update<float, float>(p, S.data(), V.data(), s, now, true, 0.0, &dd[s], &panics);
    }
    for (long s = 0; s < p.N; s++)
        for (int m = 0; m < 5; m++) dens[s * 5 + m] = dd[s].v[m];
    return GFT_OK;
}

// Per-site reference Gauss violations after step time_iter (slot (time_iter+1)%2
// holds t, the other slot t+dt), float arithmetic as in Model::calcConstraintViolation.
int oracle_gauss_site_f32(const gft_desc* d, const float* scalar, const float* vec, unsigned time_iter,
                          float* viol) {
    const Params p = make_params(d);
    const unsigned now = (time_iter + 1) % 2;
    for (long s = 0; s < p.N; s++) site_gauss<float, float>(p, scalar, vec, s, now, viol + s * p.G, nullptr);
    return GFT_OK;
}

// RandomUniform::setInitialFields (InitialConditions/Source/RandomUniform.cpp:60-77).
void oracle_ic_random_uniform(float* field, size_t n, int seed, float lo, float hi, int set_equal) {
    std::mt19937 gen(seed);
    std::uniform_real_distribution<float> dist(lo, hi);
    size_t m = set_equal ? n / 2 : n;
    for (size_t i = 0; i < m; i++) {
        field[i] = dist(gen);
        if (set_equal) field[i + m] = field[i];
    }
}

// Reference site order (for tests of the ordering logic).
long oracle_site_order(long nx, long ny, long nz, long* out) {
    std::vector<long> o = reference_site_order(nx, ny, nz);
    if (out) std::memcpy(out, o.data(), o.size() * sizeof(long));
    return static_cast<long>(o.size());
}

}  // extern "C"
