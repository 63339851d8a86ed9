// gft_host.cpp -- the C++ host API (see gft_host.hpp).
#include "gft_host.hpp"

#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>

namespace gft {

// ---- ConfigReader ---------------------------------------------------------------

ConfigReader::ConfigReader(const std::string& path) : path_(path) {
    std::ifstream ifs(path);
    if (ifs.is_open()) {
        std::stringstream ss;
        ss << ifs.rdbuf();
        text_ = ss.str();
        open_ = true;
    }
}

void ConfigReader::skip_lines(int n) {
    for (int i = 0; i < n && pos_ < text_.size(); i++) {
        const size_t e = text_.find('\n', pos_);
        pos_ = (e == std::string::npos) ? text_.size() : e + 1;
    }
}

template <class T>
T ConfigReader::value() {
    T v{};
    values(&v, 1);
    return v;
}

template <class T>
void ConfigReader::values(T* out, int n) {
    if (!open_) return;
    const size_t c = text_.find(':', pos_);
    if (c == std::string::npos) { pos_ = text_.size(); return; }
    std::istringstream is(text_.substr(c + 1));
    for (int i = 0; i < n; i++) is >> out[i];
    const std::streamoff used = is.fail() ? 0 : static_cast<std::streamoff>(is.tellg());
    pos_ = c + 1 + (used > 0 ? static_cast<size_t>(used) : 0);
}

template unsigned ConfigReader::value<unsigned>();
template int ConfigReader::value<int>();
template bool ConfigReader::value<bool>();
template double ConfigReader::value<double>();
template float ConfigReader::value<float>();
template std::string ConfigReader::value<std::string>();

static std::string cfg(const std::string& root, const std::string& lib, const std::string& name) {
    return root + "/" + lib + "/Config/" + name;
}

// ---- Potentials -----------------------------------------------------------------

SO_N::SO_N(const std::string& root, unsigned num_components) : numComponents(num_components) {
    ConfigReader r(cfg(root, "Potentials", "SO_N.cfg"));   // SO_N.cpp:12-29
    r.skip_lines(15);
    lambda = r.value<double>();
    etaSqr = r.value<double>();
}

void SO_N::fill(gft_desc& d) const {
    d.lambda = lambda;
    d.eta_sqr = etaSqr;
}

Double_SO_N::Double_SO_N(const std::string& root, unsigned total) {
    ConfigReader r(cfg(root, "Potentials", "Double_SO_N.cfg"));   // Double_SO_N.cpp:35-128
    r.skip_lines(40);
    numComponents[0] = r.value<unsigned>();
    applyPotentialNormalisation = r.value<bool>();
    potentialEnergyOffset = r.value<double>();
    if (numComponents[0] > total)
        throw std::runtime_error("POTENTIAL::DOUBLE_SO_N:: The number of field components in f1 has been set to "
                                 + std::to_string(numComponents[0]) + " and the total number has been set to "
                                 + std::to_string(total) + ".\n The total number must be greater or equal to the number in f1.");
    numComponents[1] = total - numComponents[0];
    r.skip_lines(1);
    double mm[2], ll[2];
    r.values(mm, 2);
    m1 = mm[0]; m2 = mm[1];
    r.values(ll, 2);
    l1 = ll[0]; l2 = ll[1];
    l3 = r.value<double>();
    r.skip_lines(4);
    breakToSON = r.value<bool>();
    if (breakToSON) {
        if (numComponents[0] != numComponents[1])
            throw std::runtime_error("POTENTIAL::DOUBLE_SO_N:: Attempting to use the potential parameters that explicitly "
                                     "break the symmetry down to a single SO(N) requires that the number of components in f1 "
                                     "and f2 are the same.");
        r.skip_lines(1);
        m12r = r.value<double>();
        l4p5 = r.value<double>();
        double l67[2];
        r.values(l67, 2);
        l6r = l67[0]; l7r = l67[1];
        r.skip_lines(5);
        breakToUHalfN = r.value<bool>();
        if (breakToUHalfN) {
            if (numComponents[0] % 2 != 0)
                throw std::runtime_error("POTENTIAL::DOUBLE_SO_N:: Attempting to use the potential parameters that explicitly "
                                         "break the symmetry down to a single U(N/2) requires that the number of components "
                                         "in f1 (equal to f2) is even.");
            r.skip_lines(1);
            m12i = r.value<double>();
            double a[2], b[2];
            r.values(a, 2);
            l4m5 = a[0]; l5i = a[1];
            r.values(b, 2);
            l6i = b[0]; l7i = b[1];
        }
    }
    if (numComponents[0] != 4 || numComponents[1] != 4)
        throw std::runtime_error("POTENTIAL::DOUBLE_SO_N:: the GPU evolver supports two four-component doublets");
}

void Double_SO_N::fill(gft_desc& d) const {
    d.m1 = m1; d.m2 = m2; d.l1 = l1; d.l2 = l2; d.l3 = l3;
    d.break_to_son = breakToSON; d.break_to_uhalfn = breakToUHalfN;
    d.m12r = m12r; d.l4p5 = l4p5; d.l6r = l6r; d.l7r = l7r;
    d.m12i = m12i; d.l4m5 = l4m5; d.l5i = l5i; d.l6i = l6i; d.l7i = l7i;
    d.pot_norm = applyPotentialNormalisation; d.pot_offset = potentialEnergyOffset;
}

// ---- Gradients / Wilson loops -------------------------------------------------------

namespace Gradients {
StandardModel::StandardModel(const std::string& root, unsigned ns, unsigned nv) {
    ConfigReader r(cfg(root, "Gradients", "StandardModel.cfg"));   // StandardModel.cpp:46-59
    r.skip_lines(12);
    divisionByZeroTolerance = r.value<double>();
    if (ns != 4)
        throw std::runtime_error("GRADIENTS::STANDARDMODEL:: 4 scalar components are required but " + std::to_string(ns)
                                 + " have been assigned.");
    if (nv != 12 && nv != 15)   // StandardModel.cpp:163-169
        throw std::runtime_error("GRADIENTS::STANDARDMODEL:: Either 12 or 15 vector components are required (depending "
                                 "on whether the generator or quaternion representation is used) but "
                                 + std::to_string(nv) + " have been assigned.\n"
                                 "Note that 3 spatial dimensions are assumed so requested number of components is "
                                 "multiplied by 3 internally.");
}

TwoHDM::TwoHDM(const std::string& root, unsigned ns, unsigned nv, bool generator) {
    ConfigReader r(cfg(root, "Gradients", "TwoHDM.cfg"));
    r.skip_lines(12);
    divisionByZeroTolerance = r.value<double>();
    if (ns != 8)
        throw std::runtime_error("GRADIENTS::TWOHDM:: 8 scalar components are required but " + std::to_string(ns)
                                 + " have been assigned.");
    // TwoHDM.cpp:215-242 (WilsonLoops::TwoHDM, built first, has checked the count).
    gaugeNum = generator ? (nv == 12 ? 4 : nv == 15 ? 5 : 7) : (nv == 15 ? 4 : nv == 18 ? 5 : 7);
}
}  // namespace Gradients

namespace WilsonLoops {
StandardModel::StandardModel(const std::string& root, unsigned nv) {
    ConfigReader r(cfg(root, "WilsonLoops", "StandardModel.cfg"));   // StandardModel.cpp:5-44
    r.skip_lines(12);
    g = r.value<double>();
    gp = r.value<double>();
    r.skip_lines(1);
    divisionByZeroTolerance = r.value<double>();
    if (nv != 12 && nv != 15)
        throw std::runtime_error("WILSONLOOPS::STANDARDMODEL:: Either 12 or 15 vector components are required "
                                 "(depending on whether the generator or quaternion representation is used) but "
                                 + std::to_string(nv) + " have been assigned.");
    usingGeneratorRepresentation = (nv == 12);   // StandardModel.cpp:206-209
}

float StandardModel::getSqrCouplings(unsigned c) const {
    return static_cast<float>(c % 4 == 0 ? gp * gp : g * g);
}

TwoHDM::TwoHDM(const std::string& root, unsigned nv) {
    ConfigReader r(cfg(root, "WilsonLoops", "TwoHDM.cfg"));   // TwoHDM.cpp:5-44
    r.skip_lines(12);
    g = r.value<double>();
    gp = r.value<double>();
    gpp = r.value<double>();
    r.skip_lines(1);
    usingGeneratorRepresentation = r.value<bool>();
    r.skip_lines(1);
    divisionByZeroTolerance = r.value<double>();
    // TwoHDM.cpp:215-242: the representation and the component count fix gaugeNum.
    if (usingGeneratorRepresentation ? (nv != 12 && nv != 15 && nv != 21) : (nv != 15 && nv != 18 && nv != 27))
        throw std::runtime_error(usingGeneratorRepresentation
            ? "WILSONLOOPS::TWOHDM:: When using the generator representation, either 12 (just SM), 15 (SM x U(1)) or "
              "21 (SM x SU(2)) vector components are required but " + std::to_string(nv) + " have been assigned.\n"
              "Note that 3 spatial dimensions are assumed so requested number of components is multiplied by 3 "
              "internally."
            : "WILSONLOOPS::TWOHDM:: When using the quaternion representation, either 15 (just SM), 18 (SM x U(1)) or "
              "27 (SM x SU(2)) vector components are required but " + std::to_string(nv) + " have been assigned.\n"
              "Note that 3 spatial dimensions are assumed so requested number of components is multiplied by 3 "
              "internally.");
    gaugeNum = usingGeneratorRepresentation ? (nv == 12 ? 4 : nv == 15 ? 5 : 7) : (nv == 15 ? 4 : nv == 18 ? 5 : 7);
}

float TwoHDM::getSqrCouplings(unsigned c) const {   // TwoHDM.cpp:258-267
    const unsigned m = c % gaugeNum;
    return static_cast<float>(m == 0 ? gp * gp : m / 4 == 0 ? g * g : gpp * gpp);
}
}  // namespace WilsonLoops

// ---- Model --------------------------------------------------------------------------

void Model::configure(const std::string& model_cfg, const std::string& root, unsigned ns, unsigned nv, unsigned,
                      unsigned, unsigned, double, double, double, double dt) {
    ConfigReader r(model_cfg);   // Core/Source/Model.cpp:145-234
    r.skip_lines(11);
    potentialType = r.value<int>();
    switch (potentialType) {
    case SO_N_POTENTIAL: potential = std::make_unique<SO_N>(root, ns); break;
    case DOUBLE_SO_N_POTENTIAL: potential = std::make_unique<Double_SO_N>(root, ns); break;
    default:
        throw std::runtime_error("MODEL:: The chosen potential type (" + std::to_string(potentialType) + ") is invalid.");
    }
    r.skip_lines(1);
    gradientType = r.value<int>();
    r.skip_lines(1);
    wilsonLoopType = r.value<int>();
    switch (wilsonLoopType) {
    case NULL_WILSON_LOOP: wilsonLoop = std::make_unique<NullWilsonLoop>(); break;
    case SM_WILSON_LOOP: wilsonLoop = std::make_unique<WilsonLoops::StandardModel>(root, nv); break;
    case TWOHDM_WILSON_LOOP: wilsonLoop = std::make_unique<WilsonLoops::TwoHDM>(root, nv); break;
    default:
        throw std::runtime_error("MODEL: The chosen wilson loop type (" + std::to_string(wilsonLoopType) + ") is invalid.");
    }
    switch (gradientType) {
    case GLOBAL_GRADIENT: gradient = std::make_unique<Gradients::Global>(); break;
    case SM_GRADIENT: gradient = std::make_unique<Gradients::StandardModel>(root, ns, nv); break;
    case TWOHDM_GRADIENT:
        gradient = std::make_unique<Gradients::TwoHDM>(root, ns, nv, wilsonLoop->isUsingGeneratorRepresentation());
        break;
    default:
        throw std::runtime_error("MODEL::ERROR: The chosen gradient type (" + std::to_string(gradientType) + ") is invalid.");
    }
    r.skip_lines(1);
    dampingFactor = r.value<float>();
    ntDamped = r.value<unsigned>();
    const bool override_nt = r.value<bool>();
    const float t_end = r.value<float>();
    if (override_nt) ntDamped = static_cast<int>(t_end / dt);
    const bool sm = potentialType == SO_N_POTENTIAL && gradientType == SM_GRADIENT && wilsonLoopType == SM_WILSON_LOOP;
    const bool thdm = potentialType == DOUBLE_SO_N_POTENTIAL && gradientType == TWOHDM_GRADIENT
                      && wilsonLoopType == TWOHDM_WILSON_LOOP;
    const bool global = gradientType == GLOBAL_GRADIENT && wilsonLoopType == NULL_WILSON_LOOP
                        && (potentialType == SO_N_POTENTIAL || (potentialType == DOUBLE_SO_N_POTENTIAL && ns == 8))
                        && ns >= 1 && ns <= 8;
    if (!sm && !thdm && !global)
        throw std::runtime_error("MODEL:: the GPU evolver runs the SM (potential 1, gradient 2, Wilson loop 1), the "
                                 "2HDM (potential 2, gradient 3, Wilson loop 2) or the global O(N) model (potential 1 "
                                 "or 2, gradient 1, Wilson loop 0; 1 to 8 scalar components)");
    numScalar = ns;
    getNumberOfConstraintEquations();
}

unsigned Model::getNumberOfConstraintEquations() const {
    const unsigned a = gradient->getNumberOfConstraintEquations(), b = wilsonLoop->getNumberOfConstraintEquations();
    if (a != b)   // Core/Source/Model.cpp:269-279
        throw std::runtime_error("MODEL: The number of constraint equations expected by the gradients (" + std::to_string(a)
                                 + ") does not match the number expected by the wilson loops (" + std::to_string(b) + ").");
    return b;
}

bool Model::isUsingQuaternionRepresentation() const {   // Model.cpp:263-267
    return wilsonLoop->getNumberOfEvolutionEquations() > 0 && !wilsonLoop->isUsingGeneratorRepresentation();
}

void Model::fill(gft_desc& d) const {
    d.model = gradientType == GLOBAL_GRADIENT          ? GFT_MODEL_GLOBAL
              : potentialType == DOUBLE_SO_N_POTENTIAL ? GFT_MODEL_2HDM
                                                       : GFT_MODEL_SM;
    if (d.model == GFT_MODEL_GLOBAL) {
        d.num_scalar = static_cast<int>(numScalar);
        d.potential = potentialType;
    }
    d.g = wilsonLoop->g;
    d.gp = wilsonLoop->gp;
    d.div_tol = wilsonLoop->divisionByZeroTolerance;
    d.gen_rep = wilsonLoop->isUsingGeneratorRepresentation() ? 1 : 0;
    d.gauge_num = static_cast<int>(wilsonLoop->getNumberOfConstraintEquations());
    if (auto* thdm = dynamic_cast<const WilsonLoops::TwoHDM*>(wilsonLoop.get())) d.gpp = thdm->gpp;
    potential->fill(d);
    d.damping = dampingFactor;
    d.nt_damped = ntDamped;
}

// ---- Initial conditions ---------------------------------------------------------------

RandomUniform::RandomUniform(const std::string& root) {
    ConfigReader r(cfg(root, "InitialConditions", "RandomUniform.cfg"));   // RandomUniform.cpp:12-36
    r.skip_lines(12);
    seed = r.value<int>();
    float mm[2] = {0, 0};
    r.values(mm, 2);
    minRange = mm[0]; maxRange = mm[1];
    setTimestepsEqual = r.value<bool>();
}

void RandomUniform::setInitialFields(std::vector<float>& field) const {
    std::mt19937 generator(seed);
    std::uniform_real_distribution<float> distribution(minRange, maxRange);
    const size_t m = setTimestepsEqual ? field.size() / 2 : field.size();
    for (size_t i = 0; i < m; i++) {
        field[i] = distribution(generator);
        if (setTimestepsEqual) field[i + m] = field[i];
    }
}

Gaussian::Gaussian(const std::string& root, unsigned ns) : numScalar(ns) {
    ConfigReader r(cfg(root, "InitialConditions", "Gaussian.cfg"));
    r.skip_lines(1);
    seed = r.value<unsigned>();
    sigma = r.value<double>();
    vev = r.value<double>();
}

void Gaussian::setInitialFields(std::vector<float>& field) const {
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, sigma);
    const size_t m = field.size() / 2;
    for (size_t i = 0; i < m; i++) {
        const unsigned c = i % numScalar;
        field[m + i] = static_cast<float>(nd(gen) + ((c % 4) == 2 ? vev : 0.0));
        field[i] = field[m + i];
    }
}

// ---- Analysers --------------------------------------------------------------------------

Energy::Energy(const std::string& root, const std::string& data_dir, double, double, double) : dataDir(data_dir) {
    ConfigReader r(cfg(root, "Analysers", "Energy.cfg"));   // Energy.cpp:5-57
    r.skip_lines(13);
    globalQuantitiesPath = r.value<std::string>();
    for (int i = 0; i < 6; i++) {
        globalOptions[i] = r.value<bool>();
        anyGlobalOptions = anyGlobalOptions || globalOptions[i];
    }
    globalFrequency = r.value<unsigned>();
    if (globalFrequency == 0) globalFrequency = 1;
    if (anyGlobalOptions) std::ofstream(dataDir + "/" + globalQuantitiesPath).close();
}

void Energy::timestepAnalysis(const gft_obs& row) {
    if (!row.has_energy) return;
    std::ofstream ofs(dataDir + "/" + globalQuantitiesPath, std::ios::app);   // Energy.cpp:169-189
    for (int m = 0; m < 6; m++)
        if (globalOptions[m]) ofs << row.energy[m] << " ";
    ofs << "\n";
}

GaugeCondition::GaugeCondition(const std::string& root, const std::string& data_dir) : dataDir(data_dir) {
    ConfigReader r(cfg(root, "Analysers", "GaugeCondition.cfg"));   // GaugeCondition.cpp:5-57
    r.skip_lines(13);
    globalQuantitiesPath = r.value<std::string>();
    for (int i = 0; i < 2; i++) {
        globalOptions[i] = r.value<bool>();
        anyGlobalOptions = anyGlobalOptions || globalOptions[i];
    }
    globalFrequency = r.value<unsigned>();
    if (globalFrequency == 0) globalFrequency = 1;
    if (anyGlobalOptions) std::ofstream(dataDir + "/" + globalQuantitiesPath).close();
}

void GaugeCondition::timestepAnalysis(const gft_obs& row) {
    if (!row.has_gauss) return;
    std::ofstream ofs(dataDir + "/" + globalQuantitiesPath, std::ios::app);   // GaugeCondition.cpp:167-194
    if (globalOptions[0])
        for (unsigned m = 0; m < row.num_eq; m++) ofs << static_cast<float>(row.gauss_int[m]) << " ";
    if (globalOptions[1])
        for (unsigned m = 0; m < row.num_eq; m++) ofs << static_cast<float>(row.gauss_max[m]) << " ";
    ofs << "\n";
}

OutputFields::OutputFields(const std::string& root, const std::string& data_dir, const std::vector<float>& scalar,
                           unsigned ns_, const std::vector<float>& vec, unsigned nv_)
    : dataDir(data_dir), scalarFields(scalar), vectorFields(vec), ns(ns_), nv(nv_) {
    ConfigReader r(cfg(root, "Analysers", "OutputFields.cfg"));   // OutputFields.cpp:5-53
    r.skip_lines(13);
    outputBothTimesteps = r.value<bool>();
    outputInitial = r.value<bool>();
    outputFinal = r.value<bool>();
    r.skip_lines(1);
    outputContinual = r.value<bool>();
    outputFrequency = r.value<unsigned>();
    if (outputFrequency == 0) outputFrequency = 1;
    r.skip_lines(1);
    initialAnalysisPath = r.value<std::string>();
    continualAnalysisPath = r.value<std::string>();
    finalAnalysisPath = r.value<std::string>();
}

void OutputFields::outputFields(const std::string& path, unsigned time_step) const {
    std::ofstream ofs(dataDir + "/" + path);   // OutputFields.cpp:63-133
    if (!ofs.is_open()) return;
    const size_t nsites = scalarFields.size() / (2 * ns);
    const size_t lo = outputBothTimesteps ? 0 : ((time_step + 1) % 2) * nsites;
    const size_t hi = outputBothTimesteps ? 2 * nsites : lo + nsites;
    for (size_t i = lo; i < hi; i++) {
        for (unsigned c = 0; c < ns; c++) ofs << scalarFields[i * ns + c] << " ";
        for (unsigned c = 0; c < nv; c++) ofs << vectorFields[i * nv + c] << " ";
        ofs << std::endl;
    }
}

// ---- Lattice -------------------------------------------------------------------------------

Lattice::Lattice(const std::string& config_root, const std::string& data_dir, int form_)
    : root(config_root), dataDir(data_dir), form(form_) {
    configure(cfg(root, "Core", "Lattice.cfg"));
    model.configure(cfg(root, "Core", "Model.cfg"), root, numScalarFieldComponents, numVectorFieldComponents, nx, ny,
                    nz, dx, dy, dz, dt);
    initFields();
}

Lattice::~Lattice() {
    if (ctx) gft_destroy(ctx);
}

void Lattice::initBoundaryCondition(const std::vector<int>& faces) {   // Lattice.cpp:217-355
    if (faces.size() != 6) throw std::runtime_error("LATTICE:: Exactly six face boundary conditions must be assigned.");
    fixedFaces = 0;
    for (int f = 0; f < 6; f++) {
        if (faces[f] != FIXED && faces[f] != PERIODIC)
            throw std::runtime_error("LATTICE:: Face boundary condition (" + std::to_string(f) + ") has unsupported type "
                                     + std::to_string(faces[f]) + ".");
        if (faces[f] == FIXED) fixedFaces |= 1 << f;
    }
    for (int a = 0; a < 3; a++)
        if ((faces[2 * a] == PERIODIC) != (faces[2 * a + 1] == PERIODIC))
            throw std::runtime_error("LATTICE:: Periodic boundary conditions must be assigned to both faces of an axis.");
    // Regions by how many axes they lie on the shell of (faces, then edges, then
    // corners); within that by the set of those axes in lexicographic order (x
    // before y before z), then by the sign pattern, low shell first.
    boundaryConditions.clear();
    std::vector<std::vector<int>> vectors;
    const int axis_sets[7][3] = {{0, -1, -1}, {1, -1, -1}, {2, -1, -1}, {0, 1, -1}, {0, 2, -1}, {1, 2, -1}, {0, 1, 2}};
    for (const auto& set : axis_sets) {
        const int k = set[2] >= 0 ? 3 : set[1] >= 0 ? 2 : 1;
        for (int signs = 0; signs < (1 << k); signs++) {
            std::vector<int> bv(3, 0);
            for (int j = 0; j < k; j++) bv[set[j]] = (signs >> (k - 1 - j)) & 1 ? 1 : -1;
            vectors.push_back(bv);
        }
    }
    for (const auto& bv : vectors) {
        bool fixed = false;   // a region on any fixed face is fixed
        for (int a = 0; a < 3; a++)
            if (bv[a] != 0 && faces[2 * a + (bv[a] > 0 ? 1 : 0)] == FIXED) fixed = true;
        std::unique_ptr<BoundaryCondition> bc;
        if (fixed) bc = std::make_unique<Fixed>(nx, ny, nz, bv);
        else bc = std::make_unique<Periodic>(nx, ny, nz, bv);
        if (!bc->determineResponsibilities(stencilSize)) boundaryConditions.push_back(std::move(bc));
    }
}

std::vector<std::vector<unsigned>> Lattice::determineResponsibilities(const unsigned& stencil_size) const {
    // Lattice.cpp:417-438: [s, n - s) per axis, empty when n < s.
    const unsigned n[3] = {nx, ny, nz};
    std::vector<std::vector<unsigned>> limits(3, std::vector<unsigned>(2, stencil_size));
    for (int a = 0; a < 3; a++) limits[a][1] = n[a] < stencil_size ? 0 : n[a] - stencil_size;
    return limits;
}

BoundaryCondition::BoundaryCondition(unsigned nx_, unsigned ny_, unsigned nz_, std::vector<int> bound_vector)
    : boundVector(std::move(bound_vector)), loopLimits(3, std::vector<unsigned>(2, 0)), nx(nx_), ny(ny_), nz(nz_) {
    if (boundVector.size() != 3) throw std::runtime_error("BOUNDARYCONDITION::ERROR: bound vector needs 3 components.");
}

bool BoundaryCondition::determineResponsibilities(const unsigned& s) {
    const unsigned n[3] = {nx, ny, nz};
    bool empty = false;
    for (int a = 0; a < 3; a++) {
        // An axis too short for two shells keeps only its low shell.
        const unsigned high = n[a] < 2 * s ? s : n[a] - s;
        switch (boundVector[a]) {
        case -1: loopLimits[a] = {0, s}; break;
        case 0: loopLimits[a] = {s, high}; break;
        case 1: loopLimits[a] = {high, n[a]}; break;
        default:
            throw std::runtime_error("BOUNDARYCONDITION::ERROR: Invalid boundVector component: "
                                     + std::to_string(boundVector[a]));
        }
        if (loopLimits[a][0] >= loopLimits[a][1]) empty = true;
    }
    return empty;
}

unsigned long long BoundaryCondition::numSites() const {
    unsigned long long c = 1;
    for (int a = 0; a < 3; a++) c *= loopLimits[a][1] > loopLimits[a][0] ? loopLimits[a][1] - loopLimits[a][0] : 0;
    return c;
}

void Lattice::configure(const std::string& path) {
    ConfigReader r(path);   // Core/Source/Lattice.cpp:33-170
    if (!r.is_open()) throw std::runtime_error("LATTICE:: cannot open " + path);
    r.skip_lines(10);
    unsigned n3[3];
    r.values(n3, 3);
    nx = n3[0]; ny = n3[1]; nz = n3[2];
    double d3[3];
    r.values(d3, 3);
    dx = d3[0]; dy = d3[1]; dz = d3[2];
    nt = r.value<unsigned>();
    dt = r.value<double>();
    r.skip_lines(1);
    numScalarFieldComponents = r.value<unsigned>();
    numVectorFieldComponents = 3 * r.value<unsigned>();
    r.skip_lines(2);
    if (r.value<bool>()) {
        const float crossings = r.value<float>();
        nt = static_cast<int>(crossings * static_cast<float>(nx) * dx / (2 * dt));
    }
    r.skip_lines(3);
    int ic[2] = {0, 0};
    r.values(ic, 2);
    for (int k = 0; k < 2; k++) {
        std::unique_ptr<InitialCondition>& slot = k == 0 ? scalarInitialConditions : vectorInitialConditions;
        if (ic[k] == RANDOM_UNIFORM_IC) slot = std::make_unique<RandomUniform>(root);
        else if (ic[k] == GAUSSIAN_IC && k == 0) slot = std::make_unique<Gaussian>(root, numScalarFieldComponents);
        else if (ic[k] != ZERO_IC)
            std::cout << "LATTICE::WARNING: Initial condition type (" << ic[k]
                      << ") is invalid.\nReverting to default behaviour (initialisation to zero)\n" << std::endl;
    }
    r.skip_lines(1);
    const bool want_fields = r.value<bool>();
    const bool want_energy = r.value<bool>();
    const bool want_gauss = r.value<bool>();
    r.skip_lines(7);
    std::vector<int> faces(6, 0);
    r.values(faces.data(), 6);
    initBoundaryCondition(faces);
    r.skip_lines(1);
    progressReport = r.value<bool>();
    reportFrequency = r.value<unsigned>();
    if (reportFrequency == 0) reportFrequency = 1;
    if (want_fields) {
        auto of = std::make_unique<OutputFields>(root, dataDir, scalarFields, numScalarFieldComponents, vectorFields,
                                                 numVectorFieldComponents);
        outputFields = of.get();
        analysers.push_back(std::move(of));
    }
    if (want_energy) analysers.push_back(std::make_unique<Energy>(root, dataDir, dx, dy, dz));
    if (want_gauss) analysers.push_back(std::make_unique<GaugeCondition>(root, dataDir));
}

void Lattice::initFields() {
    const size_t N = static_cast<size_t>(nx) * ny * nz;
    scalarFields.assign(2 * N * numScalarFieldComponents, 0.f);
    vectorFields.assign(2 * N * numVectorFieldComponents, 0.f);
    if (model.isUsingQuaternionRepresentation()) {   // identity links, Lattice.cpp:369-395
        const unsigned width = numVectorFieldComponents / 3;
        for (size_t i = 0; i < 2 * N; i++)
            for (unsigned d = 0; d < 3; d++) {
                vectorFields[i * numVectorFieldComponents + d * width + 1] = 1.f;
                if (width == 9) vectorFields[i * numVectorFieldComponents + d * width + 5] = 1.f;   // family SU(2)
            }
    }
    if (scalarInitialConditions) scalarInitialConditions->setInitialFields(scalarFields);
    if (vectorInitialConditions) vectorInitialConditions->setInitialFields(vectorFields);
}

void Lattice::initialAnalysis() {
    if (outputFields && outputFields->outputInitial) outputFields->outputFields(outputFields->initialAnalysisPath, 0);
}

void Lattice::saveCheckpoint(const std::string& path) {
    if (!ctx) throw std::runtime_error("checkpoint: the lattice has not been evolved on the GPU yet");
    size_t bytes = 0;
    if (gft_checkpoint_size(ctx, &bytes) != GFT_OK) throw std::runtime_error(gft_last_error(ctx));
    std::vector<char> buf(bytes);
    if (gft_checkpoint(ctx, buf.data(), bytes) != GFT_OK) throw std::runtime_error(gft_last_error(ctx));
    std::ofstream f(path, std::ios::binary);
    f.write(buf.data(), (std::streamsize)bytes);
    if (!f) throw std::runtime_error("checkpoint: cannot write " + path);
}

void Lattice::loadCheckpoint(const std::string& path) {
    gft_desc d{};
    d.form = resolvedForm();
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.dx = dx; d.dy = dy; d.dz = dz; d.dt = dt;
    d.fixed_faces = fixedFaces;
    model.fill(d);
    if (!ctx && gft_create(&d, &ctx) != GFT_OK) throw std::runtime_error(gft_last_error(nullptr));
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("checkpoint: cannot read " + path);
    const std::streamsize bytes = f.tellg();
    f.seekg(0);
    std::vector<char> buf((size_t)bytes);
    f.read(buf.data(), bytes);
    if (gft_restore(ctx, buf.data(), (size_t)bytes) != GFT_OK) throw std::runtime_error(gft_last_error(ctx));
    unsigned it = 0;
    if (gft_export_two_slot_f32(ctx, scalarFields.data(), vectorFields.data(), &it) != GFT_OK)
        throw std::runtime_error(gft_last_error(ctx));
    timeIter = it;
}

int Lattice::resolvedForm() const {
    if (form != GFT_FORM_AUTO) return form;
    gft_desc d{};
    model.fill(d);
    return d.model == GFT_MODEL_GLOBAL ? GFT_FORM_TWO_SLOT_F32 : GFT_FORM_MOMENTUM_F32;
}

void Lattice::evolve() {
    gft_desc d{};
    d.form = resolvedForm();
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.dx = dx; d.dy = dy; d.dz = dz; d.dt = dt;
    d.fixed_faces = fixedFaces;
    model.fill(d);
    if (!ctx) {
        if (gft_create(&d, &ctx) != GFT_OK) throw std::runtime_error(gft_last_error(nullptr));
        if (gft_import_two_slot_f32(ctx, scalarFields.data(), vectorFields.data(), timeIter) != GFT_OK)
            throw std::runtime_error(gft_last_error(ctx));
    }
    unsigned efreq = 0, gfreq = 0;
    for (auto& a : analysers) {
        if (a->energyFrequency()) efreq = a->energyFrequency();
        if (a->gaussFrequency()) gfreq = a->gaussFrequency();
    }
    // Chunks end where continual field dumps or progress reports fall.
    unsigned chunk = 1000;
    if (outputFields && outputFields->outputContinual) chunk = std::min(chunk, outputFields->outputFrequency);
    if (progressReport) chunk = std::min(chunk, reportFrequency);
    if (progressReport) std::cout << "Beginning evolution." << std::endl;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<gft_obs> rows(chunk);
    for (unsigned done = 0; done < nt;) {
        const unsigned n = std::min(chunk - (timeIter % chunk), nt - done);
        unsigned nrows = 0;
        // A snapshot on the device before every chunk: a link update out of the
        // asin/sqrt domain (the reference's "Panic in evolve",
        // WilsonLoops/Source/StandardModel.cpp:474-475, after which it goes on with
        // NaN) rolls the lattice back to the chunk's start, so the caller holds the
        // last valid state when the error reaches it.
        if (rollback && gft_snapshot(ctx) != GFT_OK) throw std::runtime_error(gft_last_error(ctx));
        const int rc = gft_evolve(ctx, n, efreq, gfreq, rows.data(), n, &nrows);
        if (rc == GFT_EDOMAIN && rollback) {
            const std::string msg = gft_last_error(ctx);
            unsigned back = 0;
            if (gft_rollback(ctx, &back) != GFT_OK) throw std::runtime_error(gft_last_error(ctx));
            unsigned it = 0;
            if (gft_export_two_slot_f32(ctx, scalarFields.data(), vectorFields.data(), &it) != GFT_OK)
                throw std::runtime_error(gft_last_error(ctx));
            timeIter = back;
            throw std::runtime_error(msg + " within steps " + std::to_string(back) + ".." +
                                     std::to_string(back + n - 1) + "; the lattice is rolled back to step " +
                                     std::to_string(back));
        }
        if (rc != GFT_OK) throw std::runtime_error(gft_last_error(ctx));
        for (unsigned k = 0; k < nrows; k++)
            for (auto& a : analysers) a->timestepAnalysis(rows[k]);
        done += n;
        timeIter += n;
        if (outputFields && outputFields->outputContinual && timeIter % outputFields->outputFrequency == 0) {
            unsigned it = 0;
            if (gft_export_two_slot_f32(ctx, scalarFields.data(), vectorFields.data(), &it) != GFT_OK)
                throw std::runtime_error(gft_last_error(ctx));
            outputFields->outputFields(outputFields->continualAnalysisPath + "_" + std::to_string(timeIter) + ".dat",
                                       timeIter);
        }
        if (progressReport && timeIter % reportFrequency == 0)
            std::cout << "Timestep " << timeIter << " completed.\r" << std::flush;
    }
    unsigned it = 0;
    if (gft_export_two_slot_f32(ctx, scalarFields.data(), vectorFields.data(), &it) != GFT_OK)
        throw std::runtime_error(gft_last_error(ctx));
    const std::chrono::duration<double> el = std::chrono::steady_clock::now() - t0;
    lastEvolveSeconds = el.count();
    if (progressReport) {
        std::cout << "Timestep " << nt << " completed.\n";
        std::cout << "Evolution finished in " << lastEvolveSeconds << " seconds." << std::endl;
    }
}

void Lattice::finalAnalysis() {
    if (outputFields && outputFields->outputFinal) outputFields->outputFields(outputFields->finalAnalysisPath, timeIter);
}

}  // namespace gft
