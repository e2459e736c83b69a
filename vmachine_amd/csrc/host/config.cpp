// config.cpp -- input.in parser with the reference's keys and order semantics.
#include "config.h"

#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace vmc_host {

namespace {

std::string trim(const std::string& s)
{
    const char* ws = " \t\n\r\f\v";
    const size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return "";
    const size_t e = s.find_last_not_of(ws);
    return s.substr(b, e - b + 1);
}

std::vector<std::string> split(const std::string& s)
{
    std::istringstream is(s);
    std::vector<std::string> out;
    std::string w;
    while (is >> w) out.push_back(w);
    return out;
}

// std::boolalpha reading as the reference does it: "true"/"false", anything
// else leaves the value false (the stream fails)
bool parse_bool(const std::string& s) { return s == "true"; }

int to_int(const std::string& s, const std::string& key)
{
    try {
        size_t n = 0;
        const int v = std::stoi(s, &n);
        return v;
    } catch (const std::exception&) {
        throw ConfigError("Invalid value '" + s + "' for key '" + key + "'");
    }
}

double to_double(const std::string& s, const std::string& key)
{
    try {
        return std::stod(s);
    } catch (const std::exception&) {
        throw ConfigError("Invalid value '" + s + "' for key '" + key + "'");
    }
}

// ref src/userinput.cpp:60-81 and :123-128 (one process)
void set_steps(Config& c, int steps)
{
    c.totalStepsWOEqui = steps;
    c.equilibrationSteps = int(c.totalStepsWOEqui * c.equilibrationFraction);
}

const char* element_label(const std::string& key)
{
    if (key == "gaussian") return "Gaussian";
    if (key == "slaterDeterminant") return "Slater determinant";
    if (key == "padeJastrow") return "Padé-Jastrow";
    if (key == "simpleJastrow") return "simple Jastrow";
    if (key == "RBMGaussian") return "RBM-Gaussian";
    if (key == "RBMProduct") return "RBM-product";
    if (key == "hydrogenLike") return "hydrogenlike";
    if (key == "hardCoreJastrow") return "hard-core Jastrow";
    return "";
}

std::string json_escape(const std::string& s)
{
    std::string o;
    for (char ch : s) {
        if (ch == '"' || ch == '\\') o += '\\';
        o += ch;
    }
    return o;
}

}  // namespace

std::string to_json(const Config& c)
{
    std::ostringstream os;
    auto b = [](bool v) { return v ? "true" : "false"; };
    auto q = [](const std::string& v) { return "\"" + json_escape(v) + "\""; };
    os << std::setprecision(17);
    os << "{\"numParticles\": " << c.numParticles << ", \"numDimensions\": " << c.numDimensions
       << ", \"numHiddenNodes\": " << c.numHiddenUnits << ", \"atomicNumber\": " << c.atomicNumber
       << ", \"omega\": " << c.omega << ", \"sigma\": " << c.sigma
       << ", \"stepLength\": " << c.stepLength << ", \"learningRate\": " << c.learningRate
       << ", \"maxRadius\": " << c.maxRadius << ", \"totalSpin\": " << c.totalSpin
       << ", \"numIterations\": " << c.numIterations << ", \"numSteps\": " << c.totalStepsWOEqui
       << ", \"equilibration\": " << c.equilibrationFraction
       << ", \"equilibrationSteps\": " << c.equilibrationSteps << ", \"sorting\": " << b(c.sorting)
       << ", \"hamiltonian\": " << q(c.hamiltonian) << ", \"interactionStyle\": " << q(c.interaction)
       << ", \"basis\": " << q(c.basis) << ", \"sampling\": " << q(c.sampling)
       << ", \"optimization\": " << q(c.optimization) << ", \"optGamma\": " << c.optGamma
       << ", \"optMonotonic\": " << c.optMonotonic << ", \"initialWeights\": " << q(c.initialWeights)
       << ", \"initialWeightsValue\": " << c.initialWeightsValue
       << ", \"weightsFile\": " << q(c.weightsFile) << ", \"initialState\": " << q(c.initialState)
       << ", \"elements\": [";
    for (size_t i = 0; i < c.elements.size(); i++) os << (i ? ", " : "") << q(c.elements[i]);
    os << "], \"trialWaveFunction\": " << q(trial_wave_function_label(c.elements))
       << ", \"seeded\": " << b(c.seeded) << ", \"seed\": " << c.seed
       << ", \"checkpointFreq\": " << c.checkpointFreq
       << ", \"checkConvergence\": " << b(c.checkConvergence)
       << ", \"numberOfEnergies\": " << c.numberOfEnergies << ", \"tolerance\": " << c.tolerance
       << ", \"applyAdaptiveSteps\": " << b(c.applyAdaptiveSteps)
       << ", \"rangeOfAdaptiveSteps\": " << c.rangeOfAdaptiveSteps
       << ", \"additionalSteps\": " << c.additionalSteps
       << ", \"additionalStepsLastIter\": " << c.additionalStepsLastIter
       << ", \"computeOneBodyDensity\": " << b(c.computeOneBodyDensity)
       << ", \"computeOneBodyDensity2\": " << b(c.computeOneBodyDensity2)
       << ", \"computeTwoBodyDensity\": " << b(c.computeTwoBodyDensity)
       << ", \"numberOfBins\": " << c.numberOfBins << ", \"doResampling\": " << b(c.doResampling)
       << ", \"printEnergyToFile\": " << b(c.printEnergyToFile)
       << ", \"printParametersToFile\": " << b(c.printParametersToFile)
       << ", \"path\": " << q(c.path) << "}";
    return os.str();
}

std::string trial_wave_function_label(const std::vector<std::string>& elements)
{
    // ref src/parser.cpp:340-522: a preset matches when the element labels are
    // drawn from its list, one per element, and (for lists of two or more)
    // not all equal; the first matching preset names the trial function
    struct Preset {
        const char* name;
        std::vector<std::string> labels;
    };
    static const std::vector<Preset> presets = {
        {"VMC", {"Gaussian"}},
        {"VMC", {"Gaussian", "Padé-Jastrow"}},
        {"VMC", {"Gaussian", "Padé-Jastrow", "Slater determinant"}},
        {"RBM", {"RBM-Gaussian", "RBM-product"}},
        {"RBM", {"RBM-Gaussian", "RBM-product", "Slater determinant"}},
        {"RBMPJ", {"RBM-Gaussian", "RBM-product", "Padé-Jastrow"}},
        {"RBMPJ", {"RBM-Gaussian", "RBM-product", "Padé-Jastrow", "Slater determinant"}},
        {"PRBM", {"RBM-Gaussian", "RBM-product", "partlyrestricted"}},
        {"PRBM", {"RBM-Gaussian", "RBM-product", "partlyrestricted", "Slater determinant"}},
        {"DRBM", {"RBM-Gaussian", "drbmproduct"}},
        {"DRBM", {"RBM-Gaussian", "drbmproduct", "Slater determinant"}},
        {"RBMSJ", {"RBM-Gaussian", "RBM-product", "simple Jastrow"}},
        {"RBMSJ", {"RBM-Gaussian", "RBM-product", "simple Jastrow", "Slater determinant"}},
        {"VMC", {"hydrogenlike"}},
        {"VMC", {"Slater determinant"}},
        {"VMC", {"hydrogenlike", "Padé-Jastrow"}},
        {"VMC", {"Slater determinant", "Padé-Jastrow"}},
        {"VMC", {"Slater determinant", "Gaussian"}},
        {"SSJ", {"Slater determinant", "Gaussian", "simple Jastrow"}},
        {"SSJ", {"Gaussian", "simple Jastrow"}},
        {"FNN", {"fnn"}},
        {"BVMC", {"Gaussian", "hard-core Jastrow"}},
    };
    std::vector<std::string> labels;
    std::string joined;
    for (const auto& e : elements) {
        labels.push_back(element_label(e));
        joined += "_" + labels.back();
    }
    for (const auto& p : presets) {
        if (p.labels.size() != labels.size() || labels.empty()) continue;
        bool member = true;
        for (const auto& l : labels) {
            bool found = false;
            for (const auto& q : p.labels) found |= (q == l);
            member &= found;
        }
        if (!member) continue;
        bool all_equal = true;
        for (const auto& l : labels) all_equal &= (l == labels[0]);
        if (labels.size() >= 2 && all_equal) continue;
        return p.name;
    }
    return joined;
}

Config parse_config(const std::string& file)
{
    Config c;
    set_steps(c, c.totalStepsWOEqui);
    std::ifstream in(file);
    if (!in.is_open()) throw ConfigError("File: '" + file + "': File not found");
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line.rfind("#", 0) == 0) continue;   // ref src/parser.cpp:40-45
        const size_t colon = line.find(':');
        if (colon == std::string::npos) continue;  // the reference reads no value and moves on
        const std::string key = trim(line.substr(0, colon));
        const std::vector<std::string> v = split(line.substr(colon + 1));
        if (v.empty()) {
            if (colon + 1 >= line.size()) continue;
            throw ConfigError("Key '" + key + "' has no value");
        }
        const std::string& a = v[0];
        if (key == "numParticles") {
            c.numParticles = to_int(a, key);
            c.numHiddenUnits = c.numParticles;
            c.atomicNumber = c.numParticles;
        } else if (key == "numDimensions") {
            c.numDimensions = to_int(a, key);
        } else if (key == "omega") {
            c.omega = to_double(a, key);
            c.stepLength = 0.1 / std::sqrt(c.omega);
            c.sigma = 1.0 / std::sqrt(c.omega);
        } else if (key == "atomicNumber") {
            c.atomicNumber = to_int(a, key);
        } else if (key == "learningRate") {
            c.learningRate = to_double(a, key);
        } else if (key == "maxRadius") {
            c.maxRadius = to_double(a, key);
        } else if (key == "numIterations") {
            c.numIterations = to_int(a, key);
        } else if (key == "equilibration") {
            c.equilibrationFraction = to_double(a, key);
            set_steps(c, c.totalStepsWOEqui);
        } else if (key == "numSteps") {
            set_steps(c, to_int(a, key));
        } else if (key == "numHiddenNodes") {
            c.numHiddenUnits = to_int(a, key);
        } else if (key == "sorting") {
            c.sorting = parse_bool(a);
        } else if (key == "totalSpin") {
            c.totalSpin = to_double(a, key);
        } else if (key == "stepLength") {
            c.stepLength = to_double(a, key);
        } else if (key == "checkpointFreq") {
            c.checkpointFreq = int(to_double(a, key));
        } else if (key == "checkConvergence") {
            c.checkConvergence = parse_bool(a);
        } else if (key == "applyAdaptiveSteps") {
            c.applyAdaptiveSteps = parse_bool(a);
        } else if (key == "computeOneBodyDensity") {
            c.computeOneBodyDensity = parse_bool(a);
        } else if (key == "computeOneBodyDensity2") {
            c.computeOneBodyDensity2 = parse_bool(a);
        } else if (key == "computeTwoBodyDensity") {
            c.computeTwoBodyDensity = parse_bool(a);
        } else if (key == "printEnergyToFile") {
            c.printEnergyToFile = parse_bool(a);
        } else if (key == "printParametersToFile") {
            c.printParametersToFile = parse_bool(a);
        } else if (key == "doResampling") {
            c.doResampling = parse_bool(a);
        } else if (key == "path") {
            c.path = a;
        } else if (key == "numberOfEnergies") {
            c.numberOfEnergies = to_int(a, key);
        } else if (key == "tolerance") {
            c.tolerance = to_double(a, key);
        } else if (key == "rangeOfAdaptiveSteps") {
            c.rangeOfAdaptiveSteps = to_int(a, key);
        } else if (key == "additionalSteps") {
            c.additionalSteps = to_int(a, key);
        } else if (key == "additionalStepsLastIter") {
            c.additionalStepsLastIter = to_int(a, key);
        } else if (key == "numberOfBins") {
            c.numberOfBins = to_int(a, key);
        } else if (key == "basis") {
            if (a != "hermite" && a != "hydrogenOrbital")
                throw ConfigError("Basis '" + a + "' is not implemented!");
            c.basis = a;
        } else if (key == "hamiltonian") {
            if (a != "harmonicOscillator" && a != "doubleWell" && a != "atomicNucleus" &&
                a != "ellipticalHarmonicOscillator")
                throw ConfigError("The Hamiltonian '" + a + "' does not exist!");
            c.hamiltonian = a;
        } else if (key == "optimization") {
            if (a != "adam" && a != "gd" && a != "sgd")
                throw ConfigError("Optimization method '" + a + "' does not exist!");
            c.optimization = a;
            c.optGamma = c.optMonotonic = 0;
            if ((a == "gd" || a == "sgd") && v.size() >= 3) {
                c.optGamma = to_double(v[1], key);
                c.optMonotonic = to_double(v[2], key);
            }
        } else if (key == "initialWeights") {
            if (a != "automatize" && a != "randomuniform" && a != "constant" && a != "fromfile")
                throw ConfigError("Initial parameter configuration '" + a + "' does not exist!");
            c.initialWeights = a;
            c.initialWeightsHasValue = false;
            if ((a == "randomuniform" || a == "constant") && v.size() >= 2) {
                c.initialWeightsValue = to_double(v[1], key);
                c.initialWeightsHasValue = true;
            }
            if (a == "constant" && v.size() < 2) c.initialWeightsValue = 1.0;
            if (a == "fromfile") c.weightsFile = v.size() >= 2 ? v[1] : "";
        } else if (key == "initialState") {
            if (a != "randomNormal" && a != "randomUniform")
                throw ConfigError("Initial state configuration '" + a + "' does not exist!");
            c.initialState = a;
        } else if (key == "sampling") {
            if (a != "importanceSampling" && a != "bruteForce")
                throw ConfigError("Sampling method '" + a + "' does not exist!");
            c.sampling = a;
        } else if (key == "waveFunction") {
            // ref src/parser.cpp:223-262: the preset replaces the element list;
            // "VMC" yields Slater determinant x Pade-Jastrow only
            c.elements.clear();
            if (a == "VMC") c.elements = {"slaterDeterminant", "padeJastrow"};
            else if (a == "RBM") c.elements = {"slaterDeterminant", "RBMGaussian", "RBMProduct"};
            else if (a == "RBMPJ")
                c.elements = {"slaterDeterminant", "RBMGaussian", "RBMProduct", "padeJastrow"};
            else if (a == "RBMSJ")
                c.elements = {"slaterDeterminant", "RBMGaussian", "RBMProduct", "simpleJastrow"};
            else if (a == "PRBM")
                c.elements = {"slaterDeterminant", "RBMGaussian", "RBMProduct", "partlyRestricted"};
            else throw ConfigError("Wave function '" + a + "' does not exist!");
            c.presetUsed = true;
        } else if (key == "waveFunctionElement") {
            if (!*element_label(a)) throw ConfigError("Wave function element '" + a + "' does not exist!");
            c.elements.push_back(a);
        } else if (key == "interactionStyle") {
            if (a != "noInteraction" && a != "coulomb")
                throw ConfigError("Interaction style '" + a + "' does not exist!");
            c.interaction = a;
        } else if (key == "rng") {
            if (a != "MersenneTwister")
                throw ConfigError("Random number generator '" + a + "' does not exist!");
            if (v.size() > 1) {
                c.seeded = true;
                c.seed = to_int(v[1], key);
            }
        } else {
            throw ConfigError("Invalid key '" + key + "' is passed to configuration file!");
        }
    }
    return c;
}

}  // namespace vmc_host
