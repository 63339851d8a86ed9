// gft_host.hpp -- C++ host API over the gft C ABI.
//
// The reference's class interfaces stay the host API (Core/Header/Lattice.hpp,
// Core/Header/Model.hpp, Potentials/Header/*.hpp, Gradients/Header/*.hpp,
// WilsonLoops/Header/*.hpp, Analysers/Header/*.hpp,
// InitialConditions/Header/*.hpp, BoundaryConditions/Header/*.hpp): the same
// class names (in namespace gft), the same positional .cfg files, the same
// Data/ outputs.  What changes is underneath: Lattice::evolve() hands the
// whole step loop to the GPU through gft_evolve() instead of calling the
// per-site virtual functions, so the per-site calcXXX methods of the
// reference have no counterpart here; the plugins are configuration holders
// that fill the gft_desc.
//
// Errors keep the reference's behaviour: configuration errors throw
// std::runtime_error with the reference's messages; a C-ABI failure is
// rethrown as std::runtime_error carrying gft_last_error().
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gft.h"

namespace gft {

// Lattice's default form: the fast path (GFT_FORM_MOMENTUM_F32) for the gauged
// models (SM; 2HDM with SU(2) x U(1) links or a Higgs-family U(1) or SU(2)), the
// FP32 two-slot form for the global O(N) model (whose two-slot kernel streams
// the scalars in the reference's float storage and operation order).
constexpr int GFT_FORM_AUTO = -1;

// Positional reader of the reference's .cfg format: descriptions end with ':'
// and values follow on the same line (e.g. Core/Source/Lattice.cpp:47-143).
class ConfigReader {
public:
    explicit ConfigReader(const std::string& path);
    bool is_open() const { return open_; }
    void skip_lines(int n);                 // std::getline n times
    template <class T> T value();           // getline up to ':' then >> value
    template <class T> void values(T* out, int n);
    const std::string& path() const { return path_; }
private:
    std::string path_;
    std::string text_;
    size_t pos_ = 0;
    bool open_ = false;
};

// ---- Potentials ---------------------------------------------------------------
enum Potential_Types { NULL_POTENTIAL = 0, SO_N_POTENTIAL, DOUBLE_SO_N_POTENTIAL };
enum Gradient_Types { NULL_GRADIENT = 0, GLOBAL_GRADIENT, SM_GRADIENT, TWOHDM_GRADIENT };
enum Wilson_Loop_Types { NULL_WILSON_LOOP = 0, SM_WILSON_LOOP, TWOHDM_WILSON_LOOP };
enum Initial_Condition_Types { ZERO_IC = 0, RANDOM_UNIFORM_IC, GAUSSIAN_IC };
enum Boundary_Condition_Types { UNASSIGNED_BOUNDARY_CONDITION = 0, FIXED, NEUMANN, PERIODIC };

class Potential {
public:
    virtual ~Potential() = default;
    virtual void fill(gft_desc& d) const = 0;   // potential parameters of the C ABI
};

// V = 1/4 lambda (|f|^2 - eta^2)^2 (Potentials/Source/SO_N.cpp:12-40 config, 57-87 physics).
class SO_N : public Potential {
public:
    SO_N(const std::string& config_root, unsigned num_components);
    void fill(gft_desc& d) const override;
    double lambda = 0, etaSqr = 0;
private:
    unsigned numComponents;
};

// SO(N)xSO(M) -> SO(N) -> U(N/2) potential (Potentials/Source/Double_SO_N.cpp:35-140).
class Double_SO_N : public Potential {
public:
    Double_SO_N(const std::string& config_root, unsigned total_components);
    void fill(gft_desc& d) const override;
    unsigned numComponents[2] = {0, 0};
    bool applyPotentialNormalisation = false, breakToSON = false, breakToUHalfN = false;
    double potentialEnergyOffset = 0;
    double m1 = 0, m2 = 0, m12r = 0, m12i = 0;
    double l1 = 0, l2 = 0, l3 = 0, l4p5 = 0, l4m5 = 0, l5i = 0, l6r = 0, l6i = 0, l7r = 0, l7i = 0;
};

// ---- Gradients and Wilson loops (configuration of the gauge sector) -----------
class Gradient {
public:
    virtual ~Gradient() = default;
    virtual unsigned getNumberOfConstraintEquations() const = 0;
    double divisionByZeroTolerance = 1e-9;
};

class WilsonLoop {
public:
    virtual ~WilsonLoop() = default;
    virtual bool isUsingGeneratorRepresentation() const { return usingGeneratorRepresentation; }
    virtual unsigned getNumberOfEvolutionEquations() const = 0;
    virtual unsigned getNumberOfConstraintEquations() const = 0;
    // Squared coupling of equation comp_iter, rounded through float as the
    // reference returns it (WilsonLoops/Source/StandardModel.cpp:224-230).
    virtual float getSqrCouplings(unsigned comp_iter) const = 0;
    double g = 0, gp = 0, divisionByZeroTolerance = 1e-9;
    bool usingGeneratorRepresentation = false;
};

namespace Gradients {
class Global : public Gradient {          // Gradients/Source/Global.cpp: plain Laplacian, no gauge fields
public:
    unsigned getNumberOfConstraintEquations() const override { return 0; }
};
class StandardModel : public Gradient {   // Gradients/Source/StandardModel.cpp:5-69, 152-170
public:
    StandardModel(const std::string& config_root, unsigned num_scalar, unsigned num_vector);
    unsigned getNumberOfConstraintEquations() const override { return 4; }
};
class TwoHDM : public Gradient {          // Gradients/Source/TwoHDM.cpp:200-246
public:
    TwoHDM(const std::string& config_root, unsigned num_scalar, unsigned num_vector, bool generator);
    unsigned getNumberOfConstraintEquations() const override { return gaugeNum; }
    unsigned gaugeNum = 4;
};
}  // namespace Gradients

// WilsonLoops/Source/WilsonLoop.cpp:22-85: no gauge dynamics.
class NullWilsonLoop : public WilsonLoop {
public:
    unsigned getNumberOfEvolutionEquations() const override { return 0; }
    unsigned getNumberOfConstraintEquations() const override { return 0; }
    float getSqrCouplings(unsigned) const override { return 0.f; }
};

namespace WilsonLoops {
class StandardModel : public WilsonLoop { // WilsonLoops/Source/StandardModel.cpp:5-44, 195-241
public:
    StandardModel(const std::string& config_root, unsigned num_vector);
    unsigned getNumberOfEvolutionEquations() const override { return 12; }
    unsigned getNumberOfConstraintEquations() const override { return 4; }
    float getSqrCouplings(unsigned c) const override;
};
class TwoHDM : public WilsonLoop {        // WilsonLoops/Source/TwoHDM.cpp:5-60, 215-280
public:
    TwoHDM(const std::string& config_root, unsigned num_vector);
    unsigned getNumberOfEvolutionEquations() const override { return 3 * gaugeNum; }
    unsigned getNumberOfConstraintEquations() const override { return gaugeNum; }
    float getSqrCouplings(unsigned c) const override;
    double gpp = 0;
    unsigned gaugeNum = 4;   // 4: SU(2) x U(1); 5, 7: with a Higgs-family U(1) or SU(2)
};
}  // namespace WilsonLoops

// ---- Model ---------------------------------------------------------------------
class Model {
public:
    // Core/Source/Model.cpp:145-234: reads Model.cfg, then builds the potential,
    // the Wilson loops and the gradient from their own config files.
    void configure(const std::string& model_cfg, const std::string& config_root, unsigned num_scalar,
                   unsigned num_vector, unsigned nx, unsigned ny, unsigned nz, double dx, double dy, double dz,
                   double dt);
    unsigned getNumberOfConstraintEquations() const;
    bool isUsingQuaternionRepresentation() const;
    void fill(gft_desc& d) const;             // model, couplings, potential, damping
    int potentialType = NULL_POTENTIAL, gradientType = NULL_GRADIENT, wilsonLoopType = NULL_WILSON_LOOP;
    unsigned numScalar = 0;
    float dampingFactor = 0.f;
    unsigned ntDamped = 0;
    std::unique_ptr<Potential> potential;
    std::unique_ptr<Gradient> gradient;
    std::unique_ptr<WilsonLoop> wilsonLoop;
};

// ---- Initial conditions ----------------------------------------------------------
class InitialCondition {
public:
    virtual ~InitialCondition() = default;
    virtual void setInitialFields(std::vector<float>& field) const = 0;
};

// Uniform random fields from std::mt19937 (InitialConditions/Source/RandomUniform.cpp:12-77).
class RandomUniform : public InitialCondition {
public:
    explicit RandomUniform(const std::string& config_root);
    void setInitialFields(std::vector<float>& field) const override;
    int seed = 0;
    float minRange = 0, maxRange = 0;
    bool setTimestepsEqual = false;
};

// Extension: doublet(s) at the vev (0,0,v,0) plus iid Gaussian noise of sigma per
// real component, both slots equal (the 32^3 workload's initial state).
// Config InitialConditions/Config/Gaussian.cfg: seed, sigma, vev.
class Gaussian : public InitialCondition {
public:
    Gaussian(const std::string& config_root, unsigned num_scalar);
    void setInitialFields(std::vector<float>& field) const override;
    unsigned seed = 0, numScalar = 4;
    double sigma = 0.05, vev = 1.0;
};

// ---- Analysers ---------------------------------------------------------------------
class Analyser {
public:
    virtual ~Analyser() = default;
    virtual void initialAnalysis() {}
    // Called once per measured row (energy before the update, Gauss after it).
    virtual void timestepAnalysis(const gft_obs& row) {}
    virtual void finalAnalysis() {}
    virtual unsigned energyFrequency() const { return 0; }
    virtual unsigned gaussFrequency() const { return 0; }
};

// Analysers/Source/Energy.cpp: global quantities file, one line per measured step.
class Energy : public Analyser {
public:
    Energy(const std::string& config_root, const std::string& data_dir, double dx, double dy, double dz);
    void timestepAnalysis(const gft_obs& row) override;
    unsigned energyFrequency() const override { return anyGlobalOptions ? globalFrequency : 0; }
    bool globalOptions[6] = {false, false, false, false, false, false};
    bool anyGlobalOptions = false;
    unsigned globalFrequency = 1;
    std::string globalQuantitiesPath;
private:
    std::string dataDir;
};

// Analysers/Source/GaugeCondition.cpp: integrated and max |violation| per generator.
class GaugeCondition : public Analyser {
public:
    GaugeCondition(const std::string& config_root, const std::string& data_dir);
    void timestepAnalysis(const gft_obs& row) override;
    unsigned gaussFrequency() const override { return anyGlobalOptions ? globalFrequency : 0; }
    bool globalOptions[2] = {false, false};
    bool anyGlobalOptions = false;
    unsigned globalFrequency = 1;
    std::string globalQuantitiesPath;
private:
    std::string dataDir;
};

// Analysers/Source/OutputFields.cpp: text dumps of the fields.
class OutputFields : public Analyser {
public:
    OutputFields(const std::string& config_root, const std::string& data_dir, const std::vector<float>& scalar,
                 unsigned ns, const std::vector<float>& vec, unsigned nv);
    void outputFields(const std::string& path, unsigned time_step) const;
    bool outputBothTimesteps = false, outputInitial = false, outputFinal = false, outputContinual = false;
    unsigned outputFrequency = 1;
    std::string initialAnalysisPath, continualAnalysisPath, finalAnalysisPath;
private:
    std::string dataDir;
    const std::vector<float>& scalarFields;
    const std::vector<float>& vectorFields;
    unsigned ns, nv;
};

// ---- Boundary conditions -------------------------------------------------------------
// BoundaryConditions/Header/BoundaryCondition.hpp:5-80, Fixed.hpp, Periodic.hpp.  The
// lattice's boundary regions (6 faces, 12 edges, 8 corners), each named by its bound
// vector: per axis -1 (the low shell), 0 (between the shells) or +1 (the high shell).
// A region is Fixed when any face it lies on is fixed, else Periodic
// (Lattice.cpp:292-310).  The GPU step evolves the interior and every region in one
// launch: neighbour indices wrap periodically, and sites of Fixed regions keep their
// stored values (gft_desc::fixed_faces), so these classes carry the regions' extents
// and types rather than a per-region evolve loop.
class BoundaryCondition {
public:
    BoundaryCondition(unsigned nx, unsigned ny, unsigned nz, std::vector<int> bound_vector);
    virtual ~BoundaryCondition() = default;
    // Sets loopLimits for a stencil of the given half-width; true when the region
    // holds no site (BoundaryCondition.cpp:21-70).
    bool determineResponsibilities(const unsigned& stencil_size);
    virtual int type() const = 0;   // FIXED or PERIODIC
    unsigned long long numSites() const;
    const std::vector<int> boundVector;
    std::vector<std::vector<unsigned>> loopLimits;   // [axis][begin, end)
protected:
    unsigned nx, ny, nz;
};

class Periodic : public BoundaryCondition {   // BoundaryConditions/Source/Periodic.cpp
public:
    using BoundaryCondition::BoundaryCondition;
    int type() const override { return PERIODIC; }
};

class Fixed : public BoundaryCondition {      // BoundaryConditions/Source/Fixed.cpp: sites left unchanged
public:
    using BoundaryCondition::BoundaryCondition;
    int type() const override { return FIXED; }
};

// ---- Lattice -------------------------------------------------------------------------
class Lattice {
public:
    // config_root holds <Lib>/Config/*.cfg as in the reference tree; data_dir
    // receives the outputs.  form selects the GPU state form (GFT_FORM_*):
    // GFT_FORM_TWO_SLOT_F32 keeps the reference's float storage and arithmetic
    // order; GFT_FORM_MOMENTUM_F32 is the fast path; GFT_FORM_AUTO (the default)
    // takes the fast path wherever the configured model allows it.
    Lattice(const std::string& config_root, const std::string& data_dir, int form = GFT_FORM_AUTO);
    // The form the GPU context runs: the requested one, GFT_FORM_AUTO resolved.
    int resolvedForm() const;
    // Binary checkpoint of the complete GPU state (gft_checkpoint): evolving a
    // lattice restored from it continues bitwise as the saved one would.  The
    // reference has no reader for its text field dumps (OutputFields.cpp:63-133).
    void saveCheckpoint(const std::string& path);
    void loadCheckpoint(const std::string& path);
    // Roll back to the start of the chunk on a link-update domain error before
    // reporting it (gft_snapshot / gft_rollback; one device copy of the state).
    bool rollback = true;
    ~Lattice();
    void initialAnalysis();
    void evolve();
    void finalAnalysis();

    unsigned nx = 0, ny = 0, nz = 0, nt = 0;
    double dx = 0, dy = 0, dz = 0, dt = 0;
    unsigned numScalarFieldComponents = 0, numVectorFieldComponents = 0;
    std::vector<float> scalarFields, vectorFields;
    Model model;
    bool progressReport = false;
    int fixedFaces = 0;          // fixed faces, gft_desc::fixed_faces bits
    unsigned reportFrequency = 1;
    double lastEvolveSeconds = 0;
    unsigned stencilSize = 1;    // half-width of the nearest-neighbour stencils
    // Non-empty boundary regions in the reference's order (Lattice.cpp:262-355, 706-716).
    std::vector<std::unique_ptr<BoundaryCondition>> boundaryConditions;
    // Loop limits [axis][begin, end) of the sites no boundary region holds.
    std::vector<std::vector<unsigned>> determineResponsibilities(const unsigned& stencil_size) const;

private:
    void configure(const std::string& path);
    void initBoundaryCondition(const std::vector<int>& face_boundary_condition_types);
    void initFields();
    std::string root, dataDir;
    int form;
    std::unique_ptr<InitialCondition> scalarInitialConditions, vectorInitialConditions;
    std::vector<std::unique_ptr<Analyser>> analysers;
    OutputFields* outputFields = nullptr;
    gft_ctx* ctx = nullptr;
    unsigned timeIter = 0;
};

}  // namespace gft
