// ref_harness.cpp -- drives the reference's public Model API over a periodic
// lattice to write golden vectors.  TEST INFRASTRUCTURE ONLY: built by
// oracle/Makefile (target `ref`) against the reference's Model, Potentials,
// Gradients and WilsonLoops objects, run only in the build container.
//
// The site loop restates Lattice::evolve (Core/Source/Lattice.cpp:770-853) with
// the periodic pointer stencils of BoundaryConditions/Source/Periodic.cpp:18-170:
//   scalar_pointers[axis] = {x-e_axis, x, x+e_axis}
//   vector_pointers[axis] = {x-e_axis, x, x+e_axis, x+e_axis-e_j, x+e_axis-e_k} (j<k != axis)
// all at the current slot.  Every field update, force, energy density and
// Gauss violation is computed by the reference's own compiled code.
//
// usage: gft_ref_harness MODEL_CFG NS NX NY NZ DX DY DZ DT NSTEPS EFREQ GFREQ IN OUT [NV]
//   NV : link reals per site, 15 (quaternion form, default) or 12 (generator form)
//   IN : float scalar[2][N][NS], float vector[2][N][NV]
//   OUT: per step k: float scalar[2][N][NS], float vector[2][N][NV];
//        then float dens[N][5] if k % EFREQ == 0 (potential, gradient, kinetic,
//        magnetic, electric, before the update);
//        then float viol[N][4] if k % GFREQ == 0 (after the update).
#include "Model.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

int main(int argc, char** argv) {
    if (argc != 15 && argc != 16) {
        std::fprintf(stderr, "usage: %s MODEL_CFG NS NX NY NZ DX DY DZ DT NSTEPS EFREQ GFREQ IN OUT [NV]\n", argv[0]);
        return 2;
    }
    const std::string model_cfg = argv[1];
    const unsigned Ns = std::strtoul(argv[2], nullptr, 10);
    const unsigned nx = std::strtoul(argv[3], nullptr, 10);
    const unsigned ny = std::strtoul(argv[4], nullptr, 10);
    const unsigned nz = std::strtoul(argv[5], nullptr, 10);
    const double dx = std::strtod(argv[6], nullptr);
    const double dy = std::strtod(argv[7], nullptr);
    const double dz = std::strtod(argv[8], nullptr);
    const double dt = std::strtod(argv[9], nullptr);
    const unsigned nsteps = std::strtoul(argv[10], nullptr, 10);
    const unsigned efreq = std::strtoul(argv[11], nullptr, 10);
    const unsigned gfreq = std::strtoul(argv[12], nullptr, 10);
    const unsigned Nv = argc == 16 ? std::strtoul(argv[15], nullptr, 10) : 15;
    const long long N = 1LL * nx * ny * nz;

    Model model;
    model.configure(model_cfg, Ns, Nv, nx, ny, nz, dx, dy, dz, dt, false);
    const unsigned num_eq = model.getNumberOfConstraintEquations();

    std::vector<float> S(2 * N * Ns), V(2 * N * Nv);
    FILE* fi = std::fopen(argv[13], "rb");
    if (!fi || std::fread(S.data(), sizeof(float), S.size(), fi) != S.size()
        || std::fread(V.data(), sizeof(float), V.size(), fi) != V.size()) {
        std::fprintf(stderr, "cannot read %s\n", argv[13]);
        return 1;
    }
    std::fclose(fi);
    FILE* fo = std::fopen(argv[14], "wb");
    if (!fo) return 1;

    auto idx = [&](long long x, long long y, long long z) {
        x = (x + nx) % nx; y = (y + ny) % ny; z = (z + nz) % nz;
        return (x * ny + y) * nz + z;
    };
    std::vector<float> dens(N * 5), viol(N * num_eq);

    for (unsigned k = 0; k < nsteps; k++) {
        const unsigned now = (k + 1) % 2, past = !now;
        const bool me = efreq && k % efreq == 0;
        const bool mg = gfreq && k % gfreq == 0;
        model.update(k);
        model.energyPreparation(me);
        for (int pass = 0; pass < (mg ? 2 : 1); pass++) {
            for (long long x = 0; x < nx; x++)
                for (long long y = 0; y < ny; y++)
                    for (long long z = 0; z < nz; z++) {
                        const long long s = idx(x, y, z);
                        const long long c[3] = {x, y, z};
                        auto nb = [&](int a, int d) {
                            long long q[3] = {c[0], c[1], c[2]};
                            q[a] += d;
                            return idx(q[0], q[1], q[2]);
                        };
                        std::vector<std::vector<const float*>> sp(3, std::vector<const float*>(3));
                        std::vector<std::vector<const float*>> vp(3, std::vector<const float*>(5));
                        for (int a = 0; a < 3; a++) {
                            const long long m = nb(a, -1), p = nb(a, +1);
                            sp[a][0] = &S[(now * N + m) * Ns];
                            sp[a][1] = &S[(now * N + s) * Ns];
                            sp[a][2] = &S[(now * N + p) * Ns];
                            vp[a][0] = &V[(now * N + m) * Nv];
                            vp[a][1] = &V[(now * N + s) * Nv];
                            vp[a][2] = &V[(now * N + p) * Nv];
                            int slot = 3;
                            for (int j = 0; j < 3; j++) {
                                if (j == a) continue;
                                long long q[3] = {c[0], c[1], c[2]};
                                q[a] += 1;
                                q[j] -= 1;
                                vp[a][slot++] = &V[(now * N + idx(q[0], q[1], q[2])) * Nv];
                            }
                        }
                        float* ls[2] = {&S[(past * N + s) * Ns], &S[(now * N + s) * Ns]};
                        float* lv[2] = {&V[(past * N + s) * Nv], &V[(now * N + s) * Nv]};
                        if (pass == 0) {
                            model.calcPotentialContributions(ls[1]);
                            model.calcGradientContributions(sp, vp);
                            model.calcYangMillsContributions(vp, lv);
                            if (me) {
                                dens[s * 5 + 0] = model.calcPotentialEnergy(ls[1]);
                                dens[s * 5 + 1] = model.calcGradientEnergy(sp, vp);
                                dens[s * 5 + 2] = model.calcKineticEnergy(ls);
                                dens[s * 5 + 3] = model.calcMagneticEnergy(vp);
                                dens[s * 5 + 4] = model.calcElectricEnergy(lv);
                            }
                            model.evolve(ls, lv, dt, Ns, Nv);
                        } else {
                            long long tfi = N * Nv;
                            if (now == 1) tfi = -tfi;
                            const float* cls[2] = {ls[0], ls[1]};
                            std::vector<float> v = model.calcConstraintViolation(num_eq, tfi, cls, vp);
                            for (unsigned m = 0; m < num_eq; m++) viol[s * num_eq + m] = v[m];
                        }
                    }
        }
        std::fwrite(S.data(), sizeof(float), S.size(), fo);
        std::fwrite(V.data(), sizeof(float), V.size(), fo);
        if (me) std::fwrite(dens.data(), sizeof(float), dens.size(), fo);
        if (mg) std::fwrite(viol.data(), sizeof(float), viol.size(), fo);
    }
    std::fclose(fo);
    return 0;
}
