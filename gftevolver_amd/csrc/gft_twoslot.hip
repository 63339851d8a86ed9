// gft_twoslot.hip -- the reference two-time-level leapfrog on the GPU, in the
// reference's own two-slot array-of-structs layout.
//
//   scalar[2][N][Ns], vector[2][N][3*VD]; at step k the current slot is (k+1)%2,
//   the new values overwrite slot k%2 (Core/Source/Lattice.cpp:358-367, 773-774).
//
// One thread advances one site.  Each site writes only its own old slot and
// reads the current slot of its neighbours, so the step is race-free in any
// order (Core/Source/Model.cpp:353; WilsonLoops/Source/StandardModel.cpp:471,510).
// Storage T = double is the FP64 GPU oracle; T = float keeps the reference's
// float storage.  Field arithmetic is FP64 in the reference's operation order
// (FP contraction off), so that the only differences from the CPU evolver come
// from the device's sin/cos/asin; diagnostics are FP64.  VD is the link width
// per direction: 5 for the quaternion form [y, c0, c1, c2, c3], 4 for the
// generator form [y, w1, w2, w3] (U = exp(i w.sigma)), which is converted to a
// quaternion on every read and back after the update, as the reference does.
//
// This path is the correctness anchor, not the fast path: the AoS layout does
// not coalesce (a wave touches 15 strided floats per link).  The fast path is
// gft_momentum.hip.
#include <hip/hip_runtime.h>

#include "gft_math.hpp"
#include "gft_reduce.hpp"

// FP contraction is switched off inside each function of this file (the pragma
// at the top of a compound statement is scoped to it), so the momentum-form
// kernels compiled in the same translation unit keep their FMAs.
namespace gft {

__device__ __forceinline__ long long site_shift(const Coef& c, long long s, int a, int delta) {
    long long z = s % c.nz, y = (s / c.nz) % c.ny, x = s / ((long long)c.nz * c.ny);
    if (a == 0) x = (x + delta + c.nx) % c.nx;
    else if (a == 1) y = (y + delta + c.ny) % c.ny;
    else z = (z + delta + c.nz) % c.nz;
    return (x * c.ny + y) * c.nz + z;
}

// SU(2) action on one doublet, contraction off (StandardModel.cpp:94-97).
__device__ __forceinline__ void su2_act_ref(const double* U, const double* h, double* o) {
#pragma clang fp contract(off)
    o[0] = U[0] * h[0] - U[3] * h[1] + U[2] * h[2] - U[1] * h[3];
    o[1] = U[3] * h[0] + U[0] * h[1] + U[1] * h[2] + U[2] * h[3];
    o[2] = -U[2] * h[0] - U[1] * h[1] + U[0] * h[2] + U[3] * h[3];
    o[3] = U[1] * h[0] - U[2] * h[1] - U[3] * h[2] + U[0] * h[3];
}

// getSU2Representation (WilsonLoops/Source/StandardModel.cpp:78-111,
// Gradients/Source/StandardModel.cpp:103-136).  Generator form: c0 = cos|w|,
// c = (+-) sin|w| w/|w|; at or below the division-by-zero tolerance the
// reference sets c1 = sin|w| without the conjugation sign.
template <int VD, class T>
__device__ __forceinline__ void link_rep_ref(const T* link, bool conj, double tol, double* U) {
#pragma clang fp contract(off)
    if constexpr (VD == 5) {
        link_rep(link, conj, U);
    } else {
        double wm = 0.0;
        for (int k = 1; k < 4; k++) wm += static_cast<double>(link[k]) * static_cast<double>(link[k]);
        wm = sqrt(wm);
        U[0] = cos(wm);
        U[1] = U[2] = U[3] = 0.0;
        if (wm > tol) {
            const double cf = conj ? -1.0 : 1.0;
            for (int k = 1; k < 4; k++) U[k] = cf * sin(wm) * static_cast<double>(link[k]) / wm;
        } else {
            U[1] = sin(wm);
        }
    }
}

// invertSU2Representation (WilsonLoops/Source/StandardModel.cpp:114-157):
// |w| = asin|c| (pi/2 when |c| rounds above 1), pi - |w| when c0 < 0,
// w = |w| c/|c|; below the tolerance w = (|w|, 0, 0).
template <class T>
__device__ __forceinline__ void invert_rep_ref(const double* c, double tol, T* w) {
#pragma clang fp contract(off)
    double cm = 0.0;
    for (int k = 1; k < 4; k++) cm += c[k] * c[k];
    cm = sqrt(cm);
    double ph = cm > 1 ? 1.57079632679489661923 : asin(cm);
    if (c[0] < 0) ph = 3.14159265358979323846 - ph;
    if (cm > tol) {
        for (int k = 0; k < 3; k++) w[k] = static_cast<T>(ph * c[k + 1] / cm);
    } else {
        w[0] = static_cast<T>(ph);
        w[1] = w[2] = static_cast<T>(0);
    }
}

// transform(): U(1) phase then SU(2), optionally conjugated
// (Gradients/Source/StandardModel.cpp:72-100, TwoHDM.cpp:72-104).
template <int VD, class T>
__device__ __forceinline__ void transform_ref(const double* sf, int nd, const T* link, bool conj, double tol,
                                              double* out) {
#pragma clang fp contract(off)
    const double y = static_cast<double>(link[0]);
    const double cy = cos(y);
    const double sy = sin(conj ? -y : y);
    double h[8];
    for (int c = 0; c < 4 * nd; c++) {
        const int pm = -1 + 2 * (c % 2);
        h[c] = cy * sf[c] + (pm * sy) * sf[c - pm];
    }
    double U[4];
    link_rep_ref<VD>(link, conj, tol, U);
    for (int k = 0; k < nd; k++) su2_act_ref(U, h + 4 * k, out + 4 * k);
}

// Potential derivative in the reference's operation order (SO_N.cpp:57-75,
// Double_SO_N.cpp:157-254).
__device__ void potential_derivative_ref(const Coef& d, const double* f, double* pc) {
#pragma clang fp contract(off)
    if (d.ns == 4) {
        double fsq = 0.0;
        for (int c = 0; c < 4; c++) fsq += f[c] * f[c];
        for (int c = 0; c < 4; c++) pc[c] = d.lambda * (fsq - d.eta_sqr) * f[c];
        return;
    }
    double fsq[2] = {0.0, 0.0}, dot = 0.0, anti = 0.0;
    for (int fi = 0; fi < 2; fi++) {
        for (int c = 0; c < 4; c++) fsq[fi] += f[fi * 4 + c] * f[fi * 4 + c];
    }
    if (d.break_son) {
        for (int c = 0; c < 4; c++) dot += f[c] * f[4 + c];
        if (d.break_uhalfn)
            for (int c = 0; c < 2; c++) anti += f[2 * c] * f[4 + 2 * c + 1] - f[2 * c + 1] * f[4 + 2 * c];
    }
    for (int c = 0; c < 8; c++) pc[c] = 0.0;
    for (int c = 0; c < 4; c++) {
        pc[c] += (-2.0 * d.m1 + 4.0 * d.l1 * fsq[0] + 2.0 * d.l3 * fsq[1]) * f[c];
        if (d.break_son) {
            pc[c] += (-d.m12r + 2.0 * d.l4p5 * dot + d.l6r * fsq[0] + d.l7r * fsq[1]) * f[4 + c]
                   + 2.0 * d.l6r * dot * f[c];
            if (d.break_uhalfn) {
                const int alt = 1 - 2 * (c % 2);
                pc[c] += (d.m12i + 2.0 * d.l4m5 * anti - 2.0 * d.l5i * dot - d.l6i * fsq[0] - d.l7i * fsq[1]) * alt
                             * f[4 + c + alt]
                       - 2.0 * d.l5i * anti * f[4 + c] - 2.0 * d.l6i * anti * f[c];
            }
        }
    }
    for (int c = 0; c < 4; c++) {
        const int f2 = 4 + c;
        pc[f2] += (-2.0 * d.m2 + 4.0 * d.l2 * fsq[1] + 2.0 * d.l3 * fsq[0]) * f[f2];
        if (d.break_son) {
            pc[f2] += (-d.m12r + 2.0 * d.l4p5 * dot + d.l6r * fsq[0] + d.l7r * fsq[1]) * f[c]
                    + 2.0 * d.l7r * dot * f[f2];
            if (d.break_uhalfn) {
                const int alt = 1 - 2 * (c % 2);
                pc[f2] += -(d.m12i + 2.0 * d.l4m5 * anti - 2.0 * d.l5i * dot - d.l6i * fsq[0] - d.l7i * fsq[1]) * alt
                              * f[c + alt]
                        - 2.0 * d.l5i * anti * f[c] - 2.0 * d.l7i * anti * f[f2];
            }
        }
    }
}

// One leapfrog step of every site.  measure != 0 reproduces a measured energy
// step: the last SU(2) product of the magnetic and electric terms is
// normalised (WilsonLoops/Source/StandardModel.cpp:374,433) and the five FP64
// energy densities (potential, gradient, kinetic, magnetic, electric) are
// reduced into partial[block][5].
template <class T, int VD>
__global__ void __launch_bounds__(256) twoslot_step(Coef c, T* __restrict__ S, T* __restrict__ V, unsigned now,
                                                    int measure, double damping_step, double* __restrict__ partial,
                                                    unsigned* __restrict__ err) {
#pragma clang fp contract(off)
    __shared__ double lds[4 * 5];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double dens[5] = {0, 0, 0, 0, 0};
    if (s < c.nsites && !site_frozen(c, s)) {
        const int Ns = c.ns, nd = c.nd;
        const long long N = c.nsites;
        const unsigned past = !now;
        const T* Sn = S + (long long)now * N * Ns;
        const T* Vn = V + (long long)now * N * (3 * VD);
        T* sp = S + (long long)past * N * Ns + s * Ns;
        T* vp = V + (long long)past * N * (3 * VD) + s * (3 * VD);
        const T* sn = Sn + s * Ns;
        const T* vn = Vn + s * (3 * VD);
        long long xm[3], xp[3];
        for (int a = 0; a < 3; a++) { xm[a] = site_shift(c, s, a, -1); xp[a] = site_shift(c, s, a, +1); }

        double phi[8];
        for (int k = 0; k < Ns; k++) phi[k] = static_cast<double>(sn[k]);
        double pot[8];
        potential_derivative_ref(c, phi, pot);

        double der[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int st = 0; st < 3; st++) {
            for (int a = 0; a < 3; a++) {
                double tf[8];
                if (st == 1) {
                    for (int k = 0; k < Ns; k++) tf[k] = phi[k];
                } else {
                    double f[8];
                    const T* q = Sn + (st == 0 ? xm[a] : xp[a]) * Ns;
                    for (int k = 0; k < Ns; k++) f[k] = static_cast<double>(q[k]);
                    if (st == 0) transform_ref<VD>(f, nd, Vn + xm[a] * (3 * VD) + a * VD, false, c.tol, tf);
                    else transform_ref<VD>(f, nd, vn + a * VD, true, c.tol, tf);
                }
                for (int k = 0; k < Ns; k++) der[k] += c.lap[a][st] * tf[k];
            }
        }

        double cur[12];
        for (int a = 0; a < 3; a++) {
            double tf[8], nb[8];
            transform_ref<VD>(phi, nd, vn + a * VD, false, c.tol, tf);
            const T* q = Sn + xp[a] * Ns;
            for (int k = 0; k < Ns; k++) nb[k] = static_cast<double>(q[k]);
            double j[4] = {0.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < nd; k++) {
                const double* t = tf + 4 * k;
                const double* n = nb + 4 * k;
                j[0] += 0.25 * (t[0] * n[1] - t[1] * n[0] + t[2] * n[3] - t[3] * n[2]);
                j[1] += 0.25 * (t[0] * n[3] - t[1] * n[2] + t[2] * n[1] - t[3] * n[0]);
                j[2] += 0.25 * (-t[0] * n[2] - t[1] * n[3] + t[2] * n[0] + t[3] * n[1]);
                j[3] += 0.25 * (t[0] * n[1] - t[1] * n[0] - t[2] * n[3] + t[3] * n[2]);
            }
            for (int m = 0; m < 4; m++) cur[4 * a + m] = j[m];
        }

        double mag[12];
        for (int k = 0; k < 12; k++) mag[k] = 0.0;
        double mag_e = 0.0;
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                if (j == i) continue;
                const T* Lx = vn;
                const T* Lxpi = Vn + xp[i] * (3 * VD);
                const T* Lxpj = Vn + xp[j] * (3 * VD);
                const T* Lxmj = Vn + xm[j] * (3 * VD);
                const T* Lxpimj = Vn + site_shift(c, xp[i], j, -1) * (3 * VD);
                const double th_loc = -static_cast<double>(Lxpi[j * VD]) + static_cast<double>(Lxpj[i * VD])
                                    + static_cast<double>(Lx[j * VD]) - static_cast<double>(Lx[i * VD]);
                const double th_nb = -static_cast<double>(Lxpimj[j * VD]) - static_cast<double>(Lxmj[i * VD])
                                   + static_cast<double>(Lxmj[j * VD]) + static_cast<double>(Lx[i * VD]);
                mag[4 * i] += (sin(th_loc) - sin(th_nb)) * c.inv_sp[j];
                double P1[4], P2[4], Um[4];
                link_rep_ref<VD>(Lx + i * VD, true, c.tol, P1);
                link_rep_ref<VD>(Lx + j * VD, false, c.tol, Um);
                su2_product(Um, P1, true, P1);
                link_rep_ref<VD>(Lxpj + i * VD, false, c.tol, Um);
                su2_product(Um, P1, true, P1);
                link_rep_ref<VD>(Lxpi + j * VD, true, c.tol, Um);
                su2_product(Um, P1, measure != 0, P1);
                link_rep_ref<VD>(Lxpimj + j * VD, true, c.tol, P2);
                link_rep_ref<VD>(Lxmj + i * VD, true, c.tol, Um);
                su2_product(Um, P2, true, P2);
                link_rep_ref<VD>(Lxmj + j * VD, false, c.tol, Um);
                su2_product(Um, P2, true, P2);
                link_rep_ref<VD>(Lx + i * VD, false, c.tol, Um);
                su2_product(Um, P2, false, P2);
                for (int k = 1; k < 4; k++) mag[4 * i + k] += (P1[k] - P2[k]) * c.inv_sp[j];
                if (measure)
                    mag_e += 2.0 * c.inv_sp[i] * c.inv_sp[j]
                           * (c.inv_g_sqr * (1.0 - P1[0]) + c.inv_gp_sqr * (1.0 - cos(th_loc)));
            }
        }

        double el[12];
        double el_e = 0.0;
        const T* vpast = vp;
        for (int a = 0; a < 3; a++) {
            const double la = static_cast<double>(vn[a * VD]) - static_cast<double>(vpast[a * VD]);
            el[4 * a] = sin(la);
            double Up[4], Um[4];
            link_rep_ref<VD>(vpast + a * VD, true, c.tol, Up);
            link_rep_ref<VD>(vn + a * VD, false, c.tol, Um);
            su2_product(Um, Up, measure != 0, Up);
            for (int k = 1; k < 4; k++) el[4 * a + k] = Up[k];
            if (measure) el_e += 4.0 * c.inv_sp[a] * (c.inv_g_sqr * (1.0 - Up[0]) + c.inv_gp_sqr * (1.0 - cos(la)));
        }

        if (measure) {
            dens[0] = potential_energy(c, phi);
            double ge = 0.0;
            for (int a = 0; a < 3; a++) {
                double tf[8];
                transform_ref<VD>(phi, nd, vn + a * VD, false, c.tol, tf);
                const T* q = Sn + xp[a] * Ns;
                for (int k = 0; k < Ns; k++) {
                    const double dd = -c.est[a] * tf[k] + c.est[a] * static_cast<double>(q[k]);
                    ge += dd * dd;
                }
            }
            dens[1] = 0.5 * ge;
            double ke = 0.0;
            for (int k = 0; k < Ns; k++) {
                const double dd = (static_cast<double>(sn[k]) - static_cast<double>(sp[k])) / c.dt;
                ke += dd * dd;
            }
            dens[2] = 0.5 * ke;
            dens[3] = mag_e;
            dens[4] = c.inv_sqr_dt * el_e;
        }

        // Model::evolve (Core/Source/Model.cpp:339-375).
        for (int k = 0; k < Ns; k++) {
            const double rhs = c.dt * c.dt * (der[k] - pot[k])
                             - damping_step * (static_cast<double>(sn[k]) - static_cast<double>(sp[k]));
            sp[k] = static_cast<T>(2 * static_cast<double>(sn[k]) - static_cast<double>(sp[k]) + rhs);
        }
        // WilsonLoops::StandardModel::evolve (StandardModel.cpp:456-514).
        unsigned bad = 0;
        for (int a = 0; a < 3; a++) {
            double R[4];
            for (int k = 0; k < 4; k++)
                R[k] = (1.0 - damping_step) * el[4 * a + k] + c.dt * c.dt * (mag[4 * a + k] + c.gsq[k] * cur[4 * a + k]);
            vp[a * VD] = static_cast<T>(asin(R[0]) + static_cast<double>(vn[a * VD]));
            double cms = 0.0;
            for (int k = 1; k < 4; k++) cms += R[k] * R[k];
            if (R[0] * R[0] > 1 || cms > 1) bad = 1;
            const double u[4] = {sqrt(1.0 - cms), R[1], R[2], R[3]};
            double Up[4];
            link_rep_ref<VD>(vn + a * VD, false, c.tol, Up);
            su2_product(u, Up, true, Up);
            if constexpr (VD == 4) invert_rep_ref(Up, c.tol, vp + a * VD + 1);
            else for (int k = 0; k < 4; k++) vp[a * VD + 1 + k] = static_cast<T>(Up[k]);
        }
        if (bad) atomicOr(err, 1u);
    }
    if (measure) block_reduce_store<5, 0>(dens, nullptr, lds, partial, blockIdx.x);
}

// Gauss law after a step (slot `now` = t, slot !now = t+dt), FP64
// (Core/Source/Model.cpp:323-337; Gradients/Source/StandardModel.cpp:250-270;
// WilsonLoops/Source/StandardModel.cpp:274-322).  Reduces per block:
// sum |G^a| (4), sum (sum_i |electric terms| + g_a^2 |rho^a|) (4), max |G^a| (4).
template <class T, int VD>
__global__ void __launch_bounds__(256) twoslot_gauss(Coef c, const T* __restrict__ S, const T* __restrict__ V,
                                                     unsigned now, double* __restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double lds[4 * 12];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double sums[8] = {0, 0, 0, 0, 0, 0, 0, 0}, maxs[4] = {0, 0, 0, 0};
    if (s < c.nsites && !site_frozen(c, s)) {
        const long long N = c.nsites;
        const int Ns = c.ns;
        const unsigned fut = !now;
        const T* sn = S + (long long)now * N * Ns + s * Ns;
        const T* sf = S + (long long)fut * N * Ns + s * Ns;
        const T* Vn = V + (long long)now * N * (3 * VD);
        const T* Vf = V + (long long)fut * N * (3 * VD);
        double rho[4] = {0, 0, 0, 0};
        for (int k = 0; k < c.nd; k++) {
            double t[4], n[4];
            for (int m = 0; m < 4; m++) { t[m] = static_cast<double>(sn[4 * k + m]); n[m] = static_cast<double>(sf[4 * k + m]); }
            bilinears_add(t, n, rho);
        }
        double div[4] = {0, 0, 0, 0}, mag[4] = {0, 0, 0, 0};
        for (int a = 0; a < 3; a++) {
            const long long xm = site_shift(c, s, a, -1);
            const T* Ln = Vn + s * (3 * VD) + a * VD;
            const T* Lf = Vf + s * (3 * VD) + a * VD;
            const T* Mn = Vn + xm * (3 * VD) + a * VD;
            const T* Mf = Vf + xm * (3 * VD) + a * VD;
            const double la = static_cast<double>(Lf[0]) - static_cast<double>(Ln[0]);
            const double na = static_cast<double>(Mf[0]) - static_cast<double>(Mn[0]);
            div[0] += (sin(la) - sin(na)) * c.inv_sp[a];
            mag[0] += (fabs(sin(la)) + fabs(sin(na))) * c.inv_sp[a];
            double P1[4], P2[4], Um[4];
            link_rep_ref<VD>(Lf, false, c.tol, P1);
            link_rep_ref<VD>(Ln, true, c.tol, Um);
            su2_product(Um, P1, false, P1);
            link_rep_ref<VD>(Mn, true, c.tol, P2);
            link_rep_ref<VD>(Mf, false, c.tol, Um);
            su2_product(Um, P2, false, P2);
            for (int k = 1; k < 4; k++) {
                div[k] += (P1[k] - P2[k]) * c.inv_sp[a];
                mag[k] += (fabs(P1[k]) + fabs(P2[k])) * c.inv_sp[a];
            }
        }
        for (int k = 0; k < 4; k++) {
            const double g = div[k] - c.gsq[k] * rho[k];
            sums[k] = fabs(g);
            sums[4 + k] = mag[k] + fabs(c.gsq[k] * rho[k]);
            maxs[k] = fabs(g);
        }
    }
    block_reduce_store<8, 4>(sums, maxs, lds, partial, blockIdx.x);
}

template __global__ void twoslot_step<float, 5>(Coef, float*, float*, unsigned, int, double, double*, unsigned*);
template __global__ void twoslot_step<double, 5>(Coef, double*, double*, unsigned, int, double, double*, unsigned*);
template __global__ void twoslot_step<float, 4>(Coef, float*, float*, unsigned, int, double, double*, unsigned*);
template __global__ void twoslot_step<double, 4>(Coef, double*, double*, unsigned, int, double, double*, unsigned*);
template __global__ void twoslot_gauss<float, 5>(Coef, const float*, const float*, unsigned, double*);
template __global__ void twoslot_gauss<double, 5>(Coef, const double*, const double*, unsigned, double*);
template __global__ void twoslot_gauss<float, 4>(Coef, const float*, const float*, unsigned, double*);
template __global__ void twoslot_gauss<double, 4>(Coef, const double*, const double*, unsigned, double*);

}  // namespace gft
