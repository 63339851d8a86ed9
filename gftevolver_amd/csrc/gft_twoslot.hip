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

// Periodic neighbour of site s along axis a by delta in {-1, +1}.  32-bit
// index math: the host limits a lattice to 2^28 sites.
__device__ __forceinline__ long long site_shift(const Coef& c, long long s, int a, int delta) {
    const unsigned u = static_cast<unsigned>(s), nz = c.nz, ny = c.ny, nx = c.nx;
    const unsigned t = u / nz;
    unsigned z = u - t * nz, x = t / ny, y = t - x * ny;
    if (a == 0) x = delta > 0 ? (x + 1 == nx ? 0 : x + 1) : (x == 0 ? nx - 1 : x - 1);
    else if (a == 1) y = delta > 0 ? (y + 1 == ny ? 0 : y + 1) : (y == 0 ? ny - 1 : y - 1);
    else z = delta > 0 ? (z + 1 == nz ? 0 : z + 1) : (z == 0 ? nz - 1 : z - 1);
    return static_cast<long long>((x * ny + y) * nz + z);
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
template <bool GEN, class T>
__device__ __forceinline__ void link_rep_ref(const T* link, bool conj, double tol, double* U) {
#pragma clang fp contract(off)
    if constexpr (!GEN) {
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

// Link layout per direction: [y, SU(2) (c0..c3 or w1..w3), Higgs-family part]
// with the family part at OFF: y' for gaugeNum 5, a second SU(2) for 7
// (WilsonLoops/Source/TwoHDM.cpp:215-242).
template <bool GEN, int G>
struct LinkLayout {
    static constexpr int OFF = GEN ? 4 : 5;
    static constexpr int VD = OFF + (G == 5 ? 1 : G == 7 ? OFF - 1 : 0);
};

// transform(): U(1) phase then SU(2), optionally conjugated
// (Gradients/Source/StandardModel.cpp:72-100, TwoHDM.cpp:72-104); for
// gaugeNum 5 then the Higgs-family phase, of opposite sign on the two
// doublets, for 7 the Higgs-family SU(2) on (f_i, f_i+1, f_i+4, f_i+5),
// i = 0, 2 (TwoHDM.cpp:106-141).
template <bool GEN, int G, class T>
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
    link_rep_ref<GEN>(link, conj, tol, U);
    if constexpr (G == 4) {
        for (int k = 0; k < nd; k++) su2_act_ref(U, h + 4 * k, out + 4 * k);
    } else {
        constexpr int OFF = LinkLayout<GEN, G>::OFF;
        double sm[8];
        for (int k = 0; k < 2; k++) su2_act_ref(U, h + 4 * k, sm + 4 * k);
        if constexpr (G == 5) {
            const double y2 = static_cast<double>(link[OFF]);
            const double cy2 = cos(y2);
            for (int c = 0; c < 8; c++) {
                const int pm = -1 + 2 * (c % 2);
                const int pm4 = 1 - 2 * (c / 4);
                const int sg = conj ? -pm4 : pm4;
                out[c] = cy2 * sm[c] + (pm * sin(sg * y2)) * sm[c - pm];
            }
        } else {
            double V[4];
            link_rep_ref<GEN>(link + OFF - 1, conj, tol, V);
            for (int i = 0; i < 4; i += 2) {
                out[i] = V[0] * sm[i] - V[3] * sm[i + 1] + V[2] * sm[i + 4] - V[1] * sm[i + 5];
                out[i + 1] = V[3] * sm[i] + V[0] * sm[i + 1] + V[1] * sm[i + 4] + V[2] * sm[i + 5];
                out[i + 4] = -V[2] * sm[i] - V[1] * sm[i + 1] + V[0] * sm[i + 4] + V[3] * sm[i + 5];
                out[i + 5] = V[1] * sm[i] - V[2] * sm[i + 1] - V[3] * sm[i + 4] + V[0] * sm[i + 5];
            }
        }
    }
}

// Higgs-family bilinears of the upper/lower pair i in {0, 2} of both doublets
// (TwoHDM.cpp:401-424 currents, 299-319 Gauss charges), accumulated into j[4..G-1].
template <int G>
__device__ __forceinline__ void family_bilinears_add(const double* t, const double* n, int i, double* j) {
#pragma clang fp contract(off)
    if constexpr (G == 5) {
        j[4] += 0.25 * (t[i] * n[i + 1] - t[i + 1] * n[i] - t[i + 4] * n[i + 5] + t[i + 5] * n[i + 4]);
    } else if constexpr (G == 7) {
        j[4] += 0.25 * (t[i] * n[i + 5] - t[i + 1] * n[i + 4] + t[i + 4] * n[i + 1] - t[i + 5] * n[i]);
        j[5] += 0.25 * (-t[i] * n[i + 4] - t[i + 1] * n[i + 5] + t[i + 4] * n[i] + t[i + 5] * n[i + 1]);
        j[6] += 0.25 * (t[i] * n[i + 1] - t[i + 1] * n[i] - t[i + 4] * n[i + 5] + t[i + 5] * n[i + 4]);
    }
}

// Potential derivative in the reference's operation order (SO_N.cpp:57-75,
// Double_SO_N.cpp:157-254) for d.ns components: NS = d.ns as a compile-time
// count (f and pc then stay in registers; Double_SO_N has NS = 8), or NS = 0
// to loop to d.ns at run time.
template <int NS>
__device__ __forceinline__ void potential_derivative_ref(const Coef& d, const double* f, double* pc) {
#pragma clang fp contract(off)
    const int ns = NS ? NS : d.ns;
    if (d.so_n) {
        double fsq = 0.0;
#pragma unroll
        for (int c = 0; c < ns; c++) fsq += f[c] * f[c];
#pragma unroll
        for (int c = 0; c < ns; c++) pc[c] = d.lambda * (fsq - d.eta_sqr) * f[c];
        return;
    }
    if constexpr (NS != 0 && NS != 8) return;   // Double_SO_N takes 8 components (checked by gft_create)
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
// ND: doublets (1 SM, 2 2HDM) as a compile-time count, so that the per-site
// arrays stay in registers, or 0 for c.nd at run time (the generator form and
// gaugeNum 7, whose larger bodies run faster with the loops not unrolled:
// 2.24 against 1.87 ms at 128^3 for the SM generator form).  The operation
// order is the reference's whatever ND.
template <class T, bool GEN, int G, int ND>
__global__ void __launch_bounds__(256) twoslot_step(Coef c, T* __restrict__ S, T* __restrict__ V, unsigned now,
                                                    int measure, double damping_step, double* __restrict__ partial,
                                                    unsigned* __restrict__ err) {
#pragma clang fp contract(off)
    constexpr int VD = LinkLayout<GEN, G>::VD, OFF = LinkLayout<GEN, G>::OFF;
    __shared__ double lds[4 * 5];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double dens[5] = {0, 0, 0, 0, 0};
    if (s < c.nsites && !site_frozen(c, s)) {
        static_assert(G == 4 || ND != 1, "a Higgs-family field acts on two doublets");
        const int nd = ND ? ND : c.nd, Ns = 4 * nd;
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
        potential_derivative_ref<4 * ND>(c, phi, pot);

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
                    if (st == 0) transform_ref<GEN, G>(f, nd, Vn + xm[a] * (3 * VD) + a * VD, false, c.tol, tf);
                    else transform_ref<GEN, G>(f, nd, vn + a * VD, true, c.tol, tf);
                }
                for (int k = 0; k < Ns; k++) der[k] += c.lap[a][st] * tf[k];
            }
        }

        double cur[3 * G];
        for (int a = 0; a < 3; a++) {
            double tf[8], nb[8];
            transform_ref<GEN, G>(phi, nd, vn + a * VD, false, c.tol, tf);
            const T* q = Sn + xp[a] * Ns;
            for (int k = 0; k < Ns; k++) nb[k] = static_cast<double>(q[k]);
            double j[G];
            for (int m = 0; m < G; m++) j[m] = 0.0;
            for (int k = 0; k < nd; k++) {
                const double* t = tf + 4 * k;
                const double* n = nb + 4 * k;
                j[0] += 0.25 * (t[0] * n[1] - t[1] * n[0] + t[2] * n[3] - t[3] * n[2]);
                j[1] += 0.25 * (t[0] * n[3] - t[1] * n[2] + t[2] * n[1] - t[3] * n[0]);
                j[2] += 0.25 * (-t[0] * n[2] - t[1] * n[3] + t[2] * n[0] + t[3] * n[1]);
                j[3] += 0.25 * (t[0] * n[1] - t[1] * n[0] - t[2] * n[3] + t[3] * n[2]);
                family_bilinears_add<G>(tf, nb, 2 * k, j);
            }
            for (int m = 0; m < G; m++) cur[G * a + m] = j[m];
        }

        double mag[3 * G];
        for (int k = 0; k < 3 * G; k++) mag[k] = 0.0;
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
                mag[G * i] += (sin(th_loc) - sin(th_nb)) * c.inv_sp[j];
                double P1[4], P2[4], Um[4];
                link_rep_ref<GEN>(Lx + i * VD, true, c.tol, P1);
                link_rep_ref<GEN>(Lx + j * VD, false, c.tol, Um);
                su2_product(Um, P1, true, P1);
                link_rep_ref<GEN>(Lxpj + i * VD, false, c.tol, Um);
                su2_product(Um, P1, true, P1);
                link_rep_ref<GEN>(Lxpi + j * VD, true, c.tol, Um);
                su2_product(Um, P1, measure != 0, P1);
                link_rep_ref<GEN>(Lxpimj + j * VD, true, c.tol, P2);
                link_rep_ref<GEN>(Lxmj + i * VD, true, c.tol, Um);
                su2_product(Um, P2, true, P2);
                link_rep_ref<GEN>(Lxmj + j * VD, false, c.tol, Um);
                su2_product(Um, P2, true, P2);
                link_rep_ref<GEN>(Lx + i * VD, false, c.tol, Um);
                su2_product(Um, P2, false, P2);
                for (int k = 1; k < 4; k++) mag[G * i + k] += (P1[k] - P2[k]) * c.inv_sp[j];
                if (measure)
                    mag_e += 2.0 * c.inv_sp[i] * c.inv_sp[j]
                           * (c.inv_g_sqr * (1.0 - P1[0]) + c.inv_gp_sqr * (1.0 - cos(th_loc)));
                // Higgs-family loops (WilsonLoops/Source/TwoHDM.cpp:496-552).
                if constexpr (G == 5) {
                    const double f_loc = -static_cast<double>(Lxpi[j * VD + OFF]) + static_cast<double>(Lxpj[i * VD + OFF])
                                       + static_cast<double>(Lx[j * VD + OFF]) - static_cast<double>(Lx[i * VD + OFF]);
                    const double f_nb = -static_cast<double>(Lxpimj[j * VD + OFF])
                                      - static_cast<double>(Lxmj[i * VD + OFF])
                                      + static_cast<double>(Lxmj[j * VD + OFF]) + static_cast<double>(Lx[i * VD + OFF]);
                    mag[G * i + 4] += (sin(f_loc) - sin(f_nb)) * c.inv_sp[j];
                    if (measure) mag_e += 2.0 * c.inv_sp[i] * c.inv_sp[j] * c.inv_gpp_sqr * (1.0 - cos(f_loc));
                } else if constexpr (G == 7) {
                    const int o = OFF - 1;
                    link_rep_ref<GEN>(Lx + i * VD + o, true, c.tol, P1);
                    link_rep_ref<GEN>(Lx + j * VD + o, false, c.tol, Um);
                    su2_product(Um, P1, true, P1);
                    link_rep_ref<GEN>(Lxpj + i * VD + o, false, c.tol, Um);
                    su2_product(Um, P1, true, P1);
                    link_rep_ref<GEN>(Lxpi + j * VD + o, true, c.tol, Um);
                    su2_product(Um, P1, measure != 0, P1);
                    link_rep_ref<GEN>(Lxpimj + j * VD + o, true, c.tol, P2);
                    link_rep_ref<GEN>(Lxmj + i * VD + o, true, c.tol, Um);
                    su2_product(Um, P2, true, P2);
                    link_rep_ref<GEN>(Lxmj + j * VD + o, false, c.tol, Um);
                    su2_product(Um, P2, true, P2);
                    link_rep_ref<GEN>(Lx + i * VD + o, false, c.tol, Um);
                    su2_product(Um, P2, false, P2);
                    for (int k = 1; k < 4; k++) mag[G * i + 3 + k] += (P1[k] - P2[k]) * c.inv_sp[j];
                    if (measure) mag_e += 2.0 * c.inv_sp[i] * c.inv_sp[j] * c.inv_gpp_sqr * (1.0 - P1[0]);
                }
            }
        }

        double el[3 * G];
        double el_e = 0.0;
        const T* vpast = vp;
        for (int a = 0; a < 3; a++) {
            const double la = static_cast<double>(vn[a * VD]) - static_cast<double>(vpast[a * VD]);
            el[G * a] = sin(la);
            double Up[4], Um[4];
            link_rep_ref<GEN>(vpast + a * VD, true, c.tol, Up);
            link_rep_ref<GEN>(vn + a * VD, false, c.tol, Um);
            su2_product(Um, Up, measure != 0, Up);
            for (int k = 1; k < 4; k++) el[G * a + k] = Up[k];
            if (measure) el_e += 4.0 * c.inv_sp[a] * (c.inv_g_sqr * (1.0 - Up[0]) + c.inv_gp_sqr * (1.0 - cos(la)));
            // Higgs family (WilsonLoops/Source/TwoHDM.cpp:600-634).
            if constexpr (G == 5) {
                const double lf = static_cast<double>(vn[a * VD + OFF]) - static_cast<double>(vpast[a * VD + OFF]);
                el[G * a + 4] = sin(lf);
                if (measure) el_e += 4.0 * c.inv_sp[a] * c.inv_gpp_sqr * (1.0 - cos(lf));
            } else if constexpr (G == 7) {
                link_rep_ref<GEN>(vpast + a * VD + OFF - 1, true, c.tol, Up);
                link_rep_ref<GEN>(vn + a * VD + OFF - 1, false, c.tol, Um);
                su2_product(Um, Up, measure != 0, Up);
                for (int k = 1; k < 4; k++) el[G * a + 3 + k] = Up[k];
                if (measure) el_e += 4.0 * c.inv_sp[a] * c.inv_gpp_sqr * (1.0 - Up[0]);
            }
        }

        if (measure) {
            dens[0] = potential_energy(c, phi);
            double ge = 0.0;
            for (int a = 0; a < 3; a++) {
                double tf[8];
                transform_ref<GEN, G>(phi, nd, vn + a * VD, false, c.tol, tf);
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
        // WilsonLoops::StandardModel::evolve (StandardModel.cpp:456-514), TwoHDM::evolve (TwoHDM.cpp:640-755).
        unsigned bad = 0;
        for (int a = 0; a < 3; a++) {
            double R[G];
            for (int k = 0; k < G; k++)
                R[k] = (1.0 - damping_step) * el[G * a + k] + c.dt * c.dt * (mag[G * a + k] + c.gsq[k] * cur[G * a + k]);
            vp[a * VD] = static_cast<T>(asin(R[0]) + static_cast<double>(vn[a * VD]));
            double cms = 0.0;
            for (int k = 1; k < 4; k++) cms += R[k] * R[k];
            if (R[0] * R[0] > 1 || cms > 1) bad = 1;
            const double u[4] = {sqrt(1.0 - cms), R[1], R[2], R[3]};
            double Up[4];
            link_rep_ref<GEN>(vn + a * VD, false, c.tol, Up);
            su2_product(u, Up, true, Up);
            if constexpr (GEN) invert_rep_ref(Up, c.tol, vp + a * VD + 1);
            else for (int k = 0; k < 4; k++) vp[a * VD + 1 + k] = static_cast<T>(Up[k]);
            if constexpr (G == 5) {
                vp[a * VD + OFF] = static_cast<T>(asin(R[4]) + static_cast<double>(vn[a * VD + OFF]));
                if (R[4] * R[4] > 1) bad = 1;
            } else if constexpr (G == 7) {
                double cf = 0.0;
                for (int k = 4; k < 7; k++) cf += R[k] * R[k];
                if (cf > 1) bad = 1;
                const double uf[4] = {sqrt(1.0 - cf), R[4], R[5], R[6]};
                link_rep_ref<GEN>(vn + a * VD + OFF - 1, false, c.tol, Up);
                su2_product(uf, Up, true, Up);
                if constexpr (GEN) invert_rep_ref(Up, c.tol, vp + a * VD + OFF);
                else for (int k = 0; k < 4; k++) vp[a * VD + OFF + k] = static_cast<T>(Up[k]);
            }
        }
        if (bad) atomicOr(err, 1u);
    }
    if (measure) block_reduce_store<5, 0>(dens, nullptr, lds, partial, blockIdx.x);
}

// The NS stored components of one site as doubles, in as few loads as the
// site's alignment allows (a site starts at a multiple of NS elements of an
// allocation aligned to 256 bytes): 16-byte loads when NS * sizeof(T) is a
// multiple of 16, 8-byte loads when it is a multiple of 8, else one per value.
template <class T, int NS>
__device__ __forceinline__ void load_site(const T* p, double* o) {
    constexpr int B = NS * static_cast<int>(sizeof(T));
    if constexpr (sizeof(T) == 4 && B % 16 == 0) {
#pragma unroll
        for (int v = 0; v < NS / 4; v++) {
            const float4 w = reinterpret_cast<const float4*>(p)[v];
            o[4 * v] = w.x; o[4 * v + 1] = w.y; o[4 * v + 2] = w.z; o[4 * v + 3] = w.w;
        }
    } else if constexpr (sizeof(T) == 4 && B % 8 == 0) {
#pragma unroll
        for (int v = 0; v < NS / 2; v++) {
            const float2 w = reinterpret_cast<const float2*>(p)[v];
            o[2 * v] = w.x; o[2 * v + 1] = w.y;
        }
    } else if constexpr (sizeof(T) == 8 && B % 16 == 0) {
#pragma unroll
        for (int v = 0; v < NS / 2; v++) {
            const double2 w = reinterpret_cast<const double2*>(p)[v];
            o[2 * v] = w.x; o[2 * v + 1] = w.y;
        }
    } else {
#pragma unroll
        for (int k = 0; k < NS; k++) o[k] = static_cast<double>(p[k]);
    }
}
template <class T, int NS>
__device__ __forceinline__ void store_site(const double* v, T* p) {
    constexpr int B = NS * static_cast<int>(sizeof(T));
    if constexpr (sizeof(T) == 4 && B % 16 == 0) {
#pragma unroll
        for (int w = 0; w < NS / 4; w++)
            reinterpret_cast<float4*>(p)[w] = make_float4(static_cast<float>(v[4 * w]), static_cast<float>(v[4 * w + 1]),
                                                          static_cast<float>(v[4 * w + 2]),
                                                          static_cast<float>(v[4 * w + 3]));
    } else if constexpr (sizeof(T) == 4 && B % 8 == 0) {
#pragma unroll
        for (int w = 0; w < NS / 2; w++)
            reinterpret_cast<float2*>(p)[w] = make_float2(static_cast<float>(v[2 * w]), static_cast<float>(v[2 * w + 1]));
    } else if constexpr (sizeof(T) == 8 && B % 16 == 0) {
#pragma unroll
        for (int w = 0; w < NS / 2; w++) reinterpret_cast<double2*>(p)[w] = make_double2(v[2 * w], v[2 * w + 1]);
    } else {
#pragma unroll
        for (int k = 0; k < NS; k++) p[k] = static_cast<T>(v[k]);
    }
}

// One leapfrog step of the global O(N) model (Gradients/Source/Global.cpp with
// the null Wilson loop, WilsonLoops/Source/WilsonLoop.cpp:22-85): the plain
// Laplacian in the reference's order (component, axis, stencil point;
// Global.cpp:118-135), the potential and Model::evolve's scalar leapfrog
// (Model.cpp:345-356).  measure != 0 reduces the FP64 energy densities
// (potential, gradient, kinetic; the magnetic and electric parts are zero).
// NS, the component count, is a template parameter: every loop unrolls, a
// site's components travel in vector loads and stores, and each thread decodes
// its coordinates once (two divisions) and steps to its six neighbours by
// wrapped increments.  The FP64 arithmetic is the same sequence of operations
// whatever NS and the loads, so the fields are bitwise those of the reference.
// Logical block b (sites 256 b .. 256 b + 255, energy partial b) runs on XCD k
// for b in [k nb / 8, (k + 1) nb / 8): each XCD sweeps a contiguous range of
// rows, so a row's y +- 1 and x +- 1 neighbours are mostly in its own L2 (the
// hardware deals consecutive blocks to the 8 XCDs in turn).  The grid is 8
// times the longest range; a hardware block past its XCD's range exits.
template <class T, int NS>
__global__ void __launch_bounds__(256) twoslot_step_global(Coef c, T* __restrict__ S, unsigned now, int measure,
                                                           double damping_step, double* __restrict__ partial,
                                                           unsigned nb) {
#pragma clang fp contract(off)
    __shared__ double lds[4 * 5];
    const unsigned xcd = blockIdx.x % 8, idx = blockIdx.x / 8;
    const unsigned b = static_cast<unsigned>(static_cast<unsigned long long>(xcd) * nb / 8) + idx;
    if (b >= static_cast<unsigned>(static_cast<unsigned long long>(xcd + 1) * nb / 8)) return;
    const unsigned s = b * blockDim.x + threadIdx.x;
    double dens[5] = {0, 0, 0, 0, 0};
    const unsigned N = static_cast<unsigned>(c.nsites);
    if (s < N) {
        const unsigned nz = c.nz, ny = c.ny, nx = c.nx;
        const unsigned t = s / nz, z = s - t * nz, x = t / ny, y = t - x * ny;
        if (!site_frozen(c, static_cast<int>(x), static_cast<int>(y), static_cast<int>(z))) {
            const unsigned sx = ny * nz;
            // Neighbour offsets from s, in elements of one slot's site array.
            const unsigned xm = x == 0 ? s + (nx - 1) * sx : s - sx, xp = x + 1 == nx ? s - (nx - 1) * sx : s + sx;
            const unsigned ym = y == 0 ? s + (ny - 1) * nz : s - nz, yp = y + 1 == ny ? s - (ny - 1) * nz : s + nz;
            const unsigned zm = z == 0 ? s + (nz - 1) : s - 1, zp = z + 1 == nz ? s - (nz - 1) : s + 1;
            const unsigned nb[3][3] = {{xm, s, xp}, {ym, s, yp}, {zm, s, zp}};
            const T* Sn = S + static_cast<size_t>(now) * N * NS;
            T* sp = S + static_cast<size_t>(!now) * N * NS + static_cast<size_t>(s) * NS;
            // phi and pot are 8 wide: the potential's two-doublet form reads 8.
            double phi[8] = {0, 0, 0, 0, 0, 0, 0, 0}, old[NS], f[3][3][NS], pot[8], der[NS];
            load_site<T, NS>(Sn + static_cast<size_t>(s) * NS, phi);
            load_site<T, NS>(sp, old);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                load_site<T, NS>(Sn + static_cast<size_t>(nb[a][0]) * NS, f[a][0]);
                load_site<T, NS>(Sn + static_cast<size_t>(nb[a][2]) * NS, f[a][2]);
#pragma unroll
                for (int k = 0; k < NS; k++) f[a][1][k] = phi[k];
            }
            potential_derivative_ref<NS>(c, phi, pot);
#pragma unroll
            for (int k = 0; k < NS; k++) {
                der[k] = 0.0;
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int st = 0; st < 3; st++) der[k] += c.lap[a][st] * f[a][st][k];
            }
            if (measure) {
                dens[0] = potential_energy(c, phi);
                double ge = 0.0, ke = 0.0;
#pragma unroll
                for (int k = 0; k < NS; k++)
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        const double dd = -c.est[a] * phi[k] + c.est[a] * f[a][2][k];
                        ge += dd * dd;
                    }
#pragma unroll
                for (int k = 0; k < NS; k++) {
                    const double dd = (phi[k] - old[k]) / c.dt;
                    ke += dd * dd;
                }
                dens[1] = 0.5 * ge;
                dens[2] = 0.5 * ke;
            }
            double nw[NS];
#pragma unroll
            for (int k = 0; k < NS; k++) {
                const double rhs = c.dt * c.dt * (der[k] - pot[k]) - damping_step * (phi[k] - old[k]);
                nw[k] = 2 * phi[k] - old[k] + rhs;
            }
            store_site<T, NS>(nw, sp);
        }
    }
    if (measure) block_reduce_store<5, 0>(dens, nullptr, lds, partial, b);
}

// Gauss law after a step (slot `now` = t, slot !now = t+dt), FP64
// (Core/Source/Model.cpp:323-337; Gradients/Source/StandardModel.cpp:250-270;
// WilsonLoops/Source/StandardModel.cpp:274-322; Higgs family: Gradients/Source/
// TwoHDM.cpp:299-319, WilsonLoops/Source/TwoHDM.cpp:352-386).  Reduces per
// block: sum |G^a| (G), sum (sum_i |electric terms| + g_a^2 |rho^a|) (G), max |G^a| (G).
template <class T, bool GEN, int G>
__global__ void __launch_bounds__(256) twoslot_gauss(Coef c, const T* __restrict__ S, const T* __restrict__ V,
                                                     unsigned now, double* __restrict__ partial) {
#pragma clang fp contract(off)
    constexpr int VD = LinkLayout<GEN, G>::VD, OFF = LinkLayout<GEN, G>::OFF;
    __shared__ double lds[4 * 3 * G];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double sums[2 * G], maxs[G];
    for (int k = 0; k < G; k++) sums[k] = sums[G + k] = maxs[k] = 0.0;
    if (s < c.nsites && !site_frozen(c, s)) {
        const long long N = c.nsites;
        const int Ns = c.ns;
        const unsigned fut = !now;
        const T* sn = S + (long long)now * N * Ns + s * Ns;
        const T* sf = S + (long long)fut * N * Ns + s * Ns;
        const T* Vn = V + (long long)now * N * (3 * VD);
        const T* Vf = V + (long long)fut * N * (3 * VD);
        double rho[G];
        for (int k = 0; k < G; k++) rho[k] = 0.0;
        for (int k = 0; k < c.nd; k++) {
            double t[4], n[4];
            for (int m = 0; m < 4; m++) { t[m] = static_cast<double>(sn[4 * k + m]); n[m] = static_cast<double>(sf[4 * k + m]); }
            bilinears_add(t, n, rho);
        }
        if constexpr (G > 4) {
            double t8[8], n8[8];
            for (int m = 0; m < 8; m++) { t8[m] = static_cast<double>(sn[m]); n8[m] = static_cast<double>(sf[m]); }
            for (int k = 0; k < 2; k++) family_bilinears_add<G>(t8, n8, 2 * k, rho);
        }
        double div[G], mag[G];
        for (int k = 0; k < G; k++) div[k] = mag[k] = 0.0;
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
            link_rep_ref<GEN>(Lf, false, c.tol, P1);
            link_rep_ref<GEN>(Ln, true, c.tol, Um);
            su2_product(Um, P1, false, P1);
            link_rep_ref<GEN>(Mn, true, c.tol, P2);
            link_rep_ref<GEN>(Mf, false, c.tol, Um);
            su2_product(Um, P2, false, P2);
            for (int k = 1; k < 4; k++) {
                div[k] += (P1[k] - P2[k]) * c.inv_sp[a];
                mag[k] += (fabs(P1[k]) + fabs(P2[k])) * c.inv_sp[a];
            }
            if constexpr (G == 5) {
                const double lf = static_cast<double>(Lf[OFF]) - static_cast<double>(Ln[OFF]);
                const double nf = static_cast<double>(Mf[OFF]) - static_cast<double>(Mn[OFF]);
                div[4] += (sin(lf) - sin(nf)) * c.inv_sp[a];
                mag[4] += (fabs(sin(lf)) + fabs(sin(nf))) * c.inv_sp[a];
            } else if constexpr (G == 7) {
                link_rep_ref<GEN>(Lf + OFF - 1, false, c.tol, P1);
                link_rep_ref<GEN>(Ln + OFF - 1, true, c.tol, Um);
                su2_product(Um, P1, false, P1);
                link_rep_ref<GEN>(Mn + OFF - 1, true, c.tol, P2);
                link_rep_ref<GEN>(Mf + OFF - 1, false, c.tol, Um);
                su2_product(Um, P2, false, P2);
                for (int k = 1; k < 4; k++) {
                    div[3 + k] += (P1[k] - P2[k]) * c.inv_sp[a];
                    mag[3 + k] += (fabs(P1[k]) + fabs(P2[k])) * c.inv_sp[a];
                }
            }
        }
        for (int k = 0; k < G; k++) {
            const double g = div[k] - c.gsq[k] * rho[k];
            sums[k] = fabs(g);
            sums[G + k] = mag[k] + fabs(c.gsq[k] * rho[k]);
            maxs[k] = fabs(g);
        }
    }
    block_reduce_store<2 * G, G>(sums, maxs, lds, partial, blockIdx.x);
}


}  // namespace gft
