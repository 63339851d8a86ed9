// gft_run -- drop-in for the reference program's main() (main.cpp:11-66) on the GPU.
//
//   gft_run [CONFIG_ROOT [DATA_DIR [FORM]]]
//
// CONFIG_ROOT holds the reference's config tree (<Lib>/Config/*.cfg, default
// "."); DATA_DIR receives global_quantities*.dat and field dumps (default
// CONFIG_ROOT/Data); FORM is auto (default: momentum_f32, the fast path, for
// the SM and the 2HDM with any of its gauge groups; two_slot_f32 for the global
// O(N) model), two_slot_f32 (the reference's float storage
// and operation order), two_slot_f64, momentum_f32, momentum_f32_plain or
// momentum_f64;
// FORM = dump_config parses the configuration, prints it and exits (no GPU).
#include <filesystem>
#include <iostream>
#include <string>

#include "gft_host.hpp"

int main(int argc, char** argv) {
    try {
        const std::string root = argc > 1 ? argv[1] : ".";
        const std::string data = argc > 2 ? argv[2] : root + "/Data";
        const std::string form_name = argc > 3 ? argv[3] : "auto";
        int form = gft::GFT_FORM_AUTO;
        if (form_name == "dump_config") {
            gft::Lattice lattice(root, data, form);
            gft_desc d{};
            lattice.model.fill(d);
            std::cout.precision(17);
            std::cout << "n " << lattice.nx << " " << lattice.ny << " " << lattice.nz << "\n"
                      << "d " << lattice.dx << " " << lattice.dy << " " << lattice.dz << "\n"
                      << "nt " << lattice.nt << "\ndt " << lattice.dt << "\n"
                      << "ns " << lattice.numScalarFieldComponents << "\nnv " << lattice.numVectorFieldComponents << "\n"
                      << "model " << d.model << "\ng " << d.g << "\ngp " << d.gp << "\n"
                      << "lambda " << d.lambda << "\neta_sqr " << d.eta_sqr << "\n"
                      << "m1 " << d.m1 << "\nm2 " << d.m2 << "\nl1 " << d.l1 << "\nl3 " << d.l3 << "\n"
                      << "m12r " << d.m12r << "\nl4p5 " << d.l4p5 << "\nl4m5 " << d.l4m5 << "\n"
                      << "damping " << d.damping << "\nnt_damped " << d.nt_damped << "\n"
                      << "progress " << lattice.progressReport << "\n"
                      << "fixed_faces " << lattice.fixedFaces << "\n"
                      << "gen_rep " << d.gen_rep << "\n"
                      << "gauge_num " << d.gauge_num << "\ngpp " << d.gpp << "\n"
                      << "form " << lattice.resolvedForm() << "\n";
            // Boundary regions: bound vector, type, loop limits; then the interior's limits.
            for (const auto& bc : lattice.boundaryConditions) {
                std::cout << "region";
                for (int a = 0; a < 3; a++) std::cout << " " << bc->boundVector[a];
                std::cout << " " << bc->type();
                for (int a = 0; a < 3; a++) std::cout << " " << bc->loopLimits[a][0] << " " << bc->loopLimits[a][1];
                std::cout << "\n";
            }
            const auto inner = lattice.determineResponsibilities(lattice.stencilSize);
            std::cout << "interior";
            for (int a = 0; a < 3; a++) std::cout << " " << inner[a][0] << " " << inner[a][1];
            std::cout << "\n";
            double s0 = 0;
            for (float x : lattice.scalarFields) s0 += x;
            std::cout << "scalar_sum " << s0 << "\n";
            return EXIT_SUCCESS;
        }
        if (form_name == "two_slot_f32") form = GFT_FORM_TWO_SLOT_F32;
        else if (form_name == "two_slot_f64") form = GFT_FORM_TWO_SLOT_F64;
        else if (form_name == "momentum_f32") form = GFT_FORM_MOMENTUM_F32;
        else if (form_name == "momentum_f64") form = GFT_FORM_MOMENTUM_F64;
        else if (form_name == "momentum_f32_plain") form = GFT_FORM_MOMENTUM_F32_PLAIN;
        else if (form_name != "auto") throw std::runtime_error("unknown form " + form_name);
        const std::filesystem::path dd(data);
        if (std::filesystem::exists(dd) && !std::filesystem::is_empty(dd))
            std::cerr << "WARNING: The output directory " << dd
                      << " is not empty. Existing output files with matching names may be overwritten." << std::endl;
        std::filesystem::create_directories(dd);
        std::filesystem::create_directories(dd / "Continual");
        gft::Lattice lattice(root, data, form);
        lattice.initialAnalysis();
        lattice.evolve();
        lattice.finalAnalysis();
        return EXIT_SUCCESS;
    } catch (const std::exception& e) {
        std::cerr << "Error in " << e.what() << std::endl;
        return EXIT_FAILURE;
    }
}
