// gft_math.hpp -- per-site algebra of the SU(2)xU(1)-Higgs evolver, shared by
// the two-slot kernels (reference layout) and the momentum-form kernels.
//
// Conventions (reference file:line):
//   * doublet Phi = (f0 + i f1, f2 + i f3); U(1) phase e^{+iy} on both pairs
//     (Gradients/Source/StandardModel.cpp:83-88);
//   * SU(2) link U = c0 + i c.sigma stored as the quaternion (c0, c1, c2, c3);
//     its real 4x4 action on a doublet is StandardModel.cpp:94-97;
//   * quaternion product = matrix product U_a U_b, with the reference's
//     "vector part only" and "normalise" variants
//     (WilsonLoops/Source/StandardModel.cpp:160-191);
//   * a link U_i(x) carries the fibre at x to x+i.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/gft.h"

#define GFT_HD __host__ __device__ __forceinline__

namespace gft {

// All coefficients of one model, passed by value to every kernel.
struct Coef {
    int ns, nd;                  // scalar reals per site, doublets
    int so_n;                    // SO_N potential over the ns components (else Double_SO_N)
    int nx, ny, nz;
    long long nsites;
    double dt;
    double lap[3][3];            // covariant Laplacian stencil (StandardModel.cpp:14-24)
    double est[3];               // forward-difference energy stencil 1/dx_i (StandardModel.cpp:35-42)
    double inv_sp[3];            // 1/dx_i^2 (WilsonLoops StandardModel.cpp:51-53)
    double inv_sqr_dt;
    double inv_g_sqr, inv_gp_sqr;
    double gsq[7];               // getSqrCouplings, rounded through float (StandardModel.cpp:224-230;
                                 // g''^2 for the Higgs-family generators 4..6, TwoHDM.cpp:258-267)
    double inv_gpp_sqr;          // 1/g''^2 (0 below the division-by-zero tolerance)
    double lambda, eta_sqr;      // SO_N
    double m1, m2, l1, l2, l3, m12r, l4p5, l6r, l7r, m12i, l4m5, l5i, l6i, l7i;
    int break_son, break_uhalfn, pot_norm;
    double pot_offset;
    double dvol;
    int fixed;                   // fixed faces: bit 2a low, 2a+1 high face of axis a (0: periodic)
    int G;                       // Gauss generators: 4 (SU(2) x U(1)), 5 or 7 (2HDM with the
                                 // Higgs-family U(1) or SU(2)), 0 without links
    double tol;                  // division-by-zero tolerance of the generator-form conversions
};

// A site in the one-site shell of a fixed face is frozen: not updated, left out
// of the energy and Gauss sums (Fixed.cpp:159-251; Lattice.cpp:293-312).
GFT_HD inline bool site_frozen(const Coef& c, int x, int y, int z) {
    if (!c.fixed) return false;
    return ((c.fixed & 1) && x == 0) || ((c.fixed & 2) && x == c.nx - 1) || ((c.fixed & 4) && y == 0) ||
           ((c.fixed & 8) && y == c.ny - 1) || ((c.fixed & 16) && z == 0) || ((c.fixed & 32) && z == c.nz - 1);
}
GFT_HD inline bool site_frozen(const Coef& c, long long s) {
    if (!c.fixed) return false;
    const int z = (int)(s % c.nz), y = (int)((s / c.nz) % c.ny), x = (int)(s / ((long long)c.nz * c.ny));
    return site_frozen(c, x, y, z);
}

template <class R>
GFT_HD void qmul_vec(const R* a, const R* b, R* r) {
    // Vector part of a*b (SU2Product with calc_trace = false).
    r[0] = R(0);
    r[1] = a[0] * b[1] + a[1] * b[0] - a[2] * b[3] + a[3] * b[2];
    r[2] = a[0] * b[2] + a[2] * b[0] - a[3] * b[1] + a[1] * b[3];
    r[3] = a[0] * b[3] + a[3] * b[0] - a[1] * b[2] + a[2] * b[1];
}

template <class R>
GFT_HD void qmul_full(const R* a, const R* b, R* r) {
    // Full product, no normalisation.
    R dot = a[1] * b[1] + a[2] * b[2] + a[3] * b[3];
    R t0 = a[0] * b[0] - dot;
    R t1 = a[0] * b[1] + a[1] * b[0] - a[2] * b[3] + a[3] * b[2];
    R t2 = a[0] * b[2] + a[2] * b[0] - a[3] * b[1] + a[1] * b[3];
    R t3 = a[0] * b[3] + a[3] * b[0] - a[1] * b[2] + a[2] * b[1];
    r[0] = t0; r[1] = t1; r[2] = t2; r[3] = t3;
}

// SU2Product(a, b, trace): full product normalised to unit length when trace,
// else the vector part with r[0] = 0 (WilsonLoops/Source/StandardModel.cpp:160-191).
GFT_HD void su2_product(const double* a, const double* b, bool trace, double* out) {
#pragma clang fp contract(off)
    double r[4];
    r[1] = a[0] * b[1] + a[1] * b[0] - a[2] * b[3] + a[3] * b[2];
    r[2] = a[0] * b[2] + a[2] * b[0] - a[3] * b[1] + a[1] * b[3];
    r[3] = a[0] * b[3] + a[3] * b[0] - a[1] * b[2] + a[2] * b[1];
    r[0] = 0.0;
    if (trace) {
        double dot = 0.0;
        dot += a[1] * b[1];
        dot += a[2] * b[2];
        dot += a[3] * b[3];
        r[0] = a[0] * b[0] - dot;
        double nrm = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + r[3] * r[3]);
        for (int k = 0; k < 4; k++) r[k] = r[k] / nrm;
    }
    for (int k = 0; k < 4; k++) out[k] = r[k];
}

// Quaternion of a stored link [y, c0, c1, c2, c3], optionally conjugated.
template <class T>
GFT_HD void link_rep(const T* link, bool conj, double* U) {
    U[0] = static_cast<double>(link[1]);
    for (int k = 1; k < 4; k++) U[k] = conj ? -static_cast<double>(link[1 + k]) : static_cast<double>(link[1 + k]);
}

// Real 4x4 SU(2) action on one doublet (StandardModel.cpp:94-97).
template <class R>
GFT_HD void su2_act(const R* U, const R* h, R* o) {
    o[0] = U[0] * h[0] - U[3] * h[1] + U[2] * h[2] - U[1] * h[3];
    o[1] = U[3] * h[0] + U[0] * h[1] + U[1] * h[2] + U[2] * h[3];
    o[2] = -U[2] * h[0] - U[1] * h[1] + U[0] * h[2] + U[3] * h[3];
    o[3] = U[1] * h[0] - U[2] * h[1] - U[3] * h[2] + U[0] * h[3];
}

// U(1) rotation of the pairs (0,1), (2,3) by the angle whose (cos, sin) is given.
template <class R>
GFT_HD void u1_act(R cy, R sy, const R* h, R* o) {
    o[0] = cy * h[0] - sy * h[1];
    o[1] = cy * h[1] + sy * h[0];
    o[2] = cy * h[2] - sy * h[3];
    o[3] = cy * h[3] + sy * h[2];
}

// The four bilinears of the currents / Gauss charges (StandardModel.cpp:346-358),
// accumulated into j (j += 1/4 * ...).
template <class R>
GFT_HD void bilinears_add(const R* t, const R* n, R* j) {
    j[0] += R(0.25) * (t[0] * n[1] - t[1] * n[0] + t[2] * n[3] - t[3] * n[2]);
    j[1] += R(0.25) * (t[0] * n[3] - t[1] * n[2] + t[2] * n[1] - t[3] * n[0]);
    j[2] += R(0.25) * (-t[0] * n[2] - t[1] * n[3] + t[2] * n[0] + t[3] * n[1]);
    j[3] += R(0.25) * (t[0] * n[1] - t[1] * n[0] - t[2] * n[3] + t[3] * n[2]);
}

// dV/df: SO_N (Potentials/Source/SO_N.cpp:57-75) or Double_SO_N
// (Potentials/Source/Double_SO_N.cpp:157-254), in arithmetic type R.
template <class R>
GFT_HD void potential_derivative(const Coef& c, const R* f, R* pc) {
    if (c.ns == 4) {
        R fsq = f[0] * f[0] + f[1] * f[1] + f[2] * f[2] + f[3] * f[3];
        R k = R(c.lambda) * (fsq - R(c.eta_sqr));
        for (int i = 0; i < 4; i++) pc[i] = k * f[i];
        return;
    }
    R f1 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2] + f[3] * f[3];
    R f2 = f[4] * f[4] + f[5] * f[5] + f[6] * f[6] + f[7] * f[7];
    R dot = R(0), anti = R(0);
    if (c.break_son) {
        dot = f[0] * f[4] + f[1] * f[5] + f[2] * f[6] + f[3] * f[7];
        if (c.break_uhalfn) anti = (f[0] * f[5] - f[1] * f[4]) + (f[2] * f[7] - f[3] * f[6]);
    }
    const R a1 = R(-2.0 * c.m1) + R(4.0 * c.l1) * f1 + R(2.0 * c.l3) * f2;
    const R a2 = R(-2.0 * c.m2) + R(4.0 * c.l2) * f2 + R(2.0 * c.l3) * f1;
    for (int i = 0; i < 4; i++) { pc[i] = a1 * f[i]; pc[4 + i] = a2 * f[4 + i]; }
    if (c.break_son) {
        const R b = R(-c.m12r) + R(2.0 * c.l4p5) * dot + R(c.l6r) * f1 + R(c.l7r) * f2;
        for (int i = 0; i < 4; i++) {
            pc[i] += b * f[4 + i] + R(2.0 * c.l6r) * dot * f[i];
            pc[4 + i] += b * f[i] + R(2.0 * c.l7r) * dot * f[4 + i];
        }
        if (c.break_uhalfn) {
            const R u = R(c.m12i) + R(2.0 * c.l4m5) * anti - R(2.0 * c.l5i) * dot - R(c.l6i) * f1 - R(c.l7i) * f2;
            for (int i = 0; i < 4; i++) {
                const int alt = 1 - 2 * (i % 2);
                pc[i] += u * R(alt) * f[4 + i + alt] - R(2.0 * c.l5i) * anti * f[4 + i] - R(2.0 * c.l6i) * anti * f[i];
                pc[4 + i] += -u * R(alt) * f[i + alt] - R(2.0 * c.l5i) * anti * f[i] - R(2.0 * c.l7i) * anti * f[4 + i];
            }
        }
    }
}

// Potential energy density in FP64 (SO_N.cpp:78-87; Double_SO_N.cpp:274-368).
GFT_HD double potential_energy(const Coef& c, const double* f) {
    if (c.ns == 4) {
        double fsq = f[0] * f[0] + f[1] * f[1] + f[2] * f[2] + f[3] * f[3];
        double dd = fsq - c.eta_sqr;
        return 0.25 * c.lambda * dd * dd;
    }
    if (c.so_n) {   // SO_N over any number of components (the global O(N) model)
        double fsq = 0.0;
        for (int i = 0; i < c.ns; i++) fsq += f[i] * f[i];
        double dd = fsq - c.eta_sqr;
        return 0.25 * c.lambda * dd * dd;
    }
    double f1 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2] + f[3] * f[3];
    double f2 = f[4] * f[4] + f[5] * f[5] + f[6] * f[6] + f[7] * f[7];
    double dot = 0.0, anti = 0.0;
    if (c.break_son) {
        dot = f[0] * f[4] + f[1] * f[5] + f[2] * f[6] + f[3] * f[7];
        if (c.break_uhalfn) anti = (f[0] * f[5] - f[1] * f[4]) + (f[2] * f[7] - f[3] * f[6]);
    }
    double v = -c.m1 * f1 - c.m2 * f2 + c.l1 * f1 * f1 + c.l2 * f2 * f2 + c.l3 * f1 * f2;
    if (c.break_son) {
        v += -c.m12r * dot + c.l4p5 * dot * dot + c.l6r * f1 * dot + c.l7r * f2 * dot;
        if (c.break_uhalfn)
            v += c.m12i * anti + c.l4m5 * anti * anti - 2.0 * c.l5i * dot * anti - c.l6i * f1 * anti - c.l7i * f2 * anti;
    }
    if (c.pot_norm) v += c.pot_offset;
    return v;
}

// Neumaier-compensated accumulator.
struct Kahan {
    double s, c;
    GFT_HD void init() { s = 0.0; c = 0.0; }
    GFT_HD void add(double x) {
        double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
        s = t;
    }
    GFT_HD double value() const { return s + c; }
};

}  // namespace gft
