// Test infrastructure: drives every entry point of the CPU restatement
// (gft_oracle.cpp) over every model option on small lattices, for a build with
// AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle asan).  A
// sanitizer report aborts the run with a non-zero status; the last line is
// "asan_check: N cases clean".  The arithmetic itself is checked elsewhere
// against the reference's fixtures; this only exercises the memory accesses
// and integer arithmetic of each path (odd lattice sides, the smallest sides
// the reference accepts, fixed-face shells, both link forms, every gauge group).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../include/gft.h"

extern "C" {
int oracle_evolve_f32(const gft_desc* d, float* scalar, float* vec, unsigned time_iter, unsigned nsteps,
                      unsigned energy_freq, unsigned gauss_freq, double* ref_rows, gft_obs* rows,
                      unsigned max_rows, unsigned* nrows, int nthreads, unsigned long* panics);
int oracle_evolve_f64(const gft_desc* d, double* scalar, double* vec, unsigned time_iter, unsigned nsteps,
                      unsigned energy_freq, unsigned gauss_freq, double* ref_rows, gft_obs* rows,
                      unsigned max_rows, unsigned* nrows, int nthreads, unsigned long* panics);
int oracle_energy_density_f32(const gft_desc* d, const float* scalar, const float* vec, unsigned time_iter,
                              float* dens);
int oracle_gauss_site_f32(const gft_desc* d, const float* scalar, const float* vec, unsigned time_iter,
                          float* viol);
void oracle_ic_random_uniform(float* field, size_t n, int seed, float lo, float hi, int set_equal);
long oracle_site_order(long nx, long ny, long nz, long* out);
}

namespace {

struct Case {
    std::string name;
    gft_desc d;
    int ns;      // scalar components per site
    int width;   // link reals per direction (0: no links)
};

gft_desc base(int model, int nx, int ny, int nz) {
    gft_desc d;
    std::memset(&d, 0, sizeof d);
    d.model = model;
    d.form = GFT_FORM_TWO_SLOT_F64;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.dx = d.dy = d.dz = 0.5;
    d.dt = 0.05;
    d.g = 0.65; d.gp = 0.35; d.div_tol = 1e-9;
    d.lambda = 0.13; d.eta_sqr = 1.0;
    d.gauge_num = 4; d.gpp = 0.5;
    return d;
}

void two_hdm_potential(gft_desc& d) {
    d.m1 = d.m2 = 0.25; d.l1 = d.l2 = 0.0325; d.l3 = 0.05;
    d.break_to_son = d.break_to_uhalfn = 1;
    d.m12r = 0.1; d.l4p5 = -0.0125; d.l4m5 = -0.0375;
}

std::vector<Case> cases() {
    std::vector<Case> c;
    c.push_back({"sm 5x4x6", base(GFT_MODEL_SM, 5, 4, 6), 4, 5});
    c.push_back({"sm 2x2x2", base(GFT_MODEL_SM, 2, 2, 2), 4, 5});
    {
        Case k{"sm generator form 3x5x4", base(GFT_MODEL_SM, 3, 5, 4), 4, 4};
        k.d.gen_rep = 1;
        c.push_back(k);
    }
    for (int faces : {1 | 2, 4 | 8, 16 | 32, 63}) {
        Case k{"sm fixed faces " + std::to_string(faces) + " 6x5x7", base(GFT_MODEL_SM, 6, 5, 7), 4, 5};
        k.d.fixed_faces = faces;
        c.push_back(k);
    }
    {
        Case k{"sm damped 4x4x4", base(GFT_MODEL_SM, 4, 4, 4), 4, 5};
        k.d.damping = 0.1f; k.d.nt_damped = 2;
        c.push_back(k);
    }
    for (int gen : {0, 1}) {
        for (int g : {4, 5, 7}) {
            Case k{"2hdm gauge " + std::to_string(g) + (gen ? " generator" : "") + " 4x3x5",
                   base(GFT_MODEL_2HDM, 4, 3, 5), 8, 0};
            two_hdm_potential(k.d);
            k.d.gauge_num = g;
            k.d.gen_rep = gen;
            k.width = g == 4 ? (gen ? 4 : 5) : g == 5 ? (gen ? 5 : 6) : (gen ? 7 : 9);
            c.push_back(k);
        }
    }
    {
        Case k{"global O(3) SO_N 5x3x4", base(GFT_MODEL_GLOBAL, 5, 3, 4), 3, 0};
        k.d.num_scalar = 3; k.d.potential = 1;
        c.push_back(k);
    }
    {
        Case k{"global O(8) Double_SO_N 3x4x5", base(GFT_MODEL_GLOBAL, 3, 4, 5), 8, 0};
        two_hdm_potential(k.d);
        k.d.num_scalar = 8; k.d.potential = 2;
        c.push_back(k);
    }
    return c;
}

// Near-vacuum state with links close to the identity in the case's link form;
// both time slots equal (a frozen shell needs them equal).
template <class T>
void fill(const Case& k, std::vector<T>& S, std::vector<T>& V, unsigned seed) {
    const long N = (long)k.d.nx * k.d.ny * k.d.nz;
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> u(-0.2, 0.2);
    S.assign(2 * N * k.ns, 0);
    V.assign(k.width ? 2 * N * 3 * k.width : 1, 0);
    for (long i = 0; i < N * k.ns; i++) S[i] = S[i + N * k.ns] = (T)(0.5 + u(gen));
    if (!k.width) return;
    const bool gen_form = k.d.gen_rep != 0;
    for (long s = 0; s < N; s++)
        for (int a = 0; a < 3; a++) {
            T* l = &V[(s * 3 + a) * k.width];
            auto su2 = [&](T* q) {   // a quaternion (c0..c3) or a generator (w1..w3)
                if (gen_form) {
                    for (int j = 0; j < 3; j++) q[j] = (T)u(gen);
                } else {
                    double c1 = u(gen), c2 = u(gen), c3 = u(gen);
                    q[0] = (T)std::sqrt(1.0 - c1 * c1 - c2 * c2 - c3 * c3);
                    q[1] = (T)c1; q[2] = (T)c2; q[3] = (T)c3;
                }
            };
            l[0] = (T)u(gen);                         // U(1) phase
            su2(l + 1);
            const int base_w = gen_form ? 4 : 5;
            if (k.width == base_w + 1) l[base_w] = (T)u(gen);   // family U(1)
            else if (k.width > base_w + 1) su2(l + base_w);     // family SU(2)
        }
    for (long i = 0; i < N * 3 * k.width; i++) V[i + N * 3 * k.width] = V[i];
}

template <class T>
int run_evolve(const Case& k, int nthreads) {
    std::vector<T> S, V;
    fill(k, S, V, 7);
    const unsigned nsteps = 5, max_rows = nsteps;
    std::vector<double> ref(max_rows * (6 + 2 * GFT_MAX_EQ));
    std::vector<gft_obs> rows(max_rows);
    unsigned nrows = 0;
    unsigned long panics = 0;
    int rc;
    if constexpr (sizeof(T) == 4)
        rc = oracle_evolve_f32(&k.d, S.data(), V.data(), 0, nsteps, 1, 2, ref.data(), rows.data(), max_rows, &nrows,
                               nthreads, &panics);
    else
        rc = oracle_evolve_f64(&k.d, S.data(), V.data(), 0, nsteps, 1, 2, ref.data(), rows.data(), max_rows, &nrows,
                               nthreads, &panics);
    if (rc != GFT_OK || nrows != nsteps) {
        std::fprintf(stderr, "asan_check: %s: evolve returned %d with %u rows\n", k.name.c_str(), rc, nrows);
        return 1;
    }
    return 0;
}

int run_case(const Case& k) {
    int bad = run_evolve<float>(k, 1) + run_evolve<double>(k, 1) + run_evolve<float>(k, 3);
    const long N = (long)k.d.nx * k.d.ny * k.d.nz;
    std::vector<float> S, V;
    fill(k, S, V, 11);
    std::vector<float> dens(N * 5);
    if (oracle_energy_density_f32(&k.d, S.data(), V.data(), 2, dens.data()) != GFT_OK) bad++;
    if (k.d.model != GFT_MODEL_GLOBAL) {
        std::vector<float> viol(N * GFT_MAX_EQ);
        if (oracle_gauss_site_f32(&k.d, S.data(), V.data(), 3, viol.data()) != GFT_OK) bad++;
    }
    return bad;
}

}  // namespace

int main() {
    int bad = 0, n = 0;
    for (const Case& k : cases()) {
        bad += run_case(k);
        n++;
    }
    std::vector<float> f(2 * 37);
    oracle_ic_random_uniform(f.data(), f.size(), 42, -1.f, 1.f, 1);
    oracle_ic_random_uniform(f.data(), f.size() - 1, 43, -1.f, 1.f, 0);
    for (long s : {2L, 3L, 7L}) {
        std::vector<long> order(s * (s + 1) * (s + 2));
        if (oracle_site_order(s, s + 1, s + 2, order.data()) != (long)order.size()) bad++;
    }
    if (bad) {
        std::fprintf(stderr, "asan_check: %d failures\n", bad);
        return 1;
    }
    std::printf("asan_check: %d cases clean\n", n);
    return 0;
}
