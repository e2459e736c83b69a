// config.h -- the reference's input.in configuration, as parsed by the host.
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace vmc_host {

// A configuration error: the reference aborts with exit status 143
// (ref src/parser.cpp:126, 143, 166, 195, 208, 221, 260, 285, 298, 312, 319).
struct ConfigError : std::runtime_error {
    explicit ConfigError(const std::string& m) : std::runtime_error(m) {}
};

struct Config {
    // system (ref src/main.cpp:8-10 preamble, src/system.h:266-310 defaults)
    int numParticles = 2;
    int numDimensions = 2;
    int numHiddenUnits = 2;       // numParticles resets it (ref src/parser.cpp:51-54)
    int atomicNumber = 2;
    double omega = 1.0;
    double sigma = 1.0;           // omega sets it to 1/sqrt(omega) (ref src/parser.cpp:57-60)
    double stepLength = 0.05;
    double learningRate = 0.1;
    double maxRadius = 5;
    double totalSpin = 0;
    bool sorting = true;
    std::string hamiltonian;      // none by default (ref src/system.h:315)
    std::string interaction = "noInteraction";
    std::string basis = "none";
    std::string initialState = "randomNormal";
    std::string sampling = "importanceSampling";
    std::string optimization = "adam";
    double optGamma = 0, optMonotonic = 0;     // sgd / gd arguments
    std::string initialWeights = "automatize";
    double initialWeightsValue = 1.0;          // constant / randomuniform argument
    bool initialWeightsHasValue = false;
    std::string weightsFile;                   // fromfile argument
    std::vector<std::string> elements;         // element keys, in parameter-row order
    bool presetUsed = false;
    // random numbers
    bool seeded = false;
    int64_t seed = 0;
    // schedule (ref src/userinput.cpp:60-81, src/system.h:266-286)
    int numIterations = 1000;
    double equilibrationFraction = 0.001;      // set by the reference's main
    int totalStepsWOEqui = 1 << 19;
    int equilibrationSteps = 0;
    int checkpointFreq = 100;
    bool checkConvergence = false;
    int numberOfEnergies = 5;
    double tolerance = 1e-7;
    bool applyAdaptiveSteps = false;
    int rangeOfAdaptiveSteps = 10;
    int additionalSteps = 4;
    int additionalStepsLastIter = 8;
    // output
    bool computeOneBodyDensity = false, computeOneBodyDensity2 = false,
         computeTwoBodyDensity = false;
    bool printEnergyToFile = false, printParametersToFile = false, doResampling = false;
    std::string path;
    int numberOfBins = 1000;
};

// Parses an input.in file with the reference's key set and order semantics
// (ref src/parser.cpp:25-333).  Throws ConfigError on an unknown key or value.
Config parse_config(const std::string& file);

// The resolved configuration as one JSON object (keys as in input.in where
// the reference has one).
std::string to_json(const Config& c);

// The reference's trial-wave-function label for an element list, used by
// initialWeights: automatize (ref src/parser.cpp:396-522).
std::string trial_wave_function_label(const std::vector<std::string>& elements);

}  // namespace vmc_host
