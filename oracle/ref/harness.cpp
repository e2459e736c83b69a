// Golden-vector harness for the reference build in oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY.  It links the reference's own objects (compiled by
// oracle/Makefile from the sources under the reference tree, never copied) and
// drives the reference's System, wave-function elements, Metropolis samplers,
// Sampler and optimisers with scripted random numbers, so that the numbers it
// prints are the reference's answers for inputs chosen by the test generator
// (tests/golden/make_golden.py).  Nothing here runs with the GPU tests.
//
// usage: vmc_ref_harness <input.in> <script> <out.json>
//   script lines (whitespace separated):
//     mode elements|iterations
//     sampling importance|brute
//     steps <n>            elements mode: number of single-step snapshots
//     iterations <n>       iterations mode: optimisation iterations
//     positions <ND values>
//     ints <count> <values...>      nextInt results (coordinate indices)
//     gauss <count> <values...>     nextGaussian standard normals
//     unif <count> <values...>      nextDouble uniforms
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "system.h"
#include "sampler.h"
#include "Hamiltonians/hamiltonian.h"
#include "Interaction/interaction.h"
#include "Metropolis/bruteforce.h"
#include "Metropolis/importancesampling.h"
#include "Optimization/optimization.h"
#include "RNG/rng.h"
#include "WaveFunctions/wavefunction.h"

namespace {

class ScriptedRNG : public RandomNumberGenerator {
public:
    std::vector<int> ints;
    std::vector<double> gauss, unif;
    Eigen::VectorXd positions;
    size_t ii = 0, gi = 0, ui = 0;

    void setSeed(int) override {}
    int nextInt(int upper) override {
        int v = ints.at(ii++);
        if (v < 0 || v >= upper) {
            std::fprintf(stderr, "scripted int %d out of range %d\n", v, upper);
            std::exit(3);
        }
        return v;
    }
    double nextDouble(double low, double high) override { return low + (high - low) * unif.at(ui++); }
    double nextGaussian(double mean, double sd) override { return mean + sd * gauss.at(gi++); }
    std::string getLabel() override { return "scripted"; }
    Eigen::MatrixXd randomUniformMatrix(Eigen::Index r, Eigen::Index c, double, double) override {
        std::fprintf(stderr, "randomUniformMatrix(%ld,%ld) not scripted: use fromfile weights\n",
                     (long)r, (long)c);
        std::exit(3);
    }
    Eigen::MatrixXd randomNormalMatrix(Eigen::Index r, Eigen::Index c, double, double) override {
        if (r * c != positions.size()) {
            std::fprintf(stderr, "positions size mismatch\n");
            std::exit(3);
        }
        Eigen::MatrixXd m(r, c);
        for (Eigen::Index i = 0; i < r * c; i++) m(i) = positions(i);
        return m;
    }
};

struct Out {
    std::ostringstream s;
    bool first = true;
    void num(double v) {
        char buf[64];
        std::snprintf(buf, sizeof(buf), "%.17g", v);
        s << buf;
    }
    void vec(const Eigen::VectorXd& v) {
        s << "[";
        for (Eigen::Index i = 0; i < v.size(); i++) {
            if (i) s << ",";
            num(v(i));
        }
        s << "]";
    }
    void mat(const Eigen::MatrixXd& m) {
        s << "[";
        for (Eigen::Index i = 0; i < m.rows(); i++) {
            if (i) s << ",";
            vec(m.row(i).transpose());
        }
        s << "]";
    }
};

void dump_state(Out& o, System* sys, Metropolis* met, int step, bool with_ratio)
{
    std::vector<WaveFunction*> els = sys->getWaveFunctionElements();
    int nd = sys->getNumberOfFreeDimensions();
    o.s << "{\"step\":" << step << ",\"positions\":";
    o.vec(met->updatePositions());
    o.s << ",\"elements\":[";
    for (size_t e = 0; e < els.size(); e++) {
        if (e) o.s << ",";
        Eigen::VectorXd g(nd);
        for (int k = 0; k < nd; k++) g(k) = els[e]->computeGradient(k);
        o.s << "{\"label\":\"" << e << "\",";
        if (with_ratio) {
            o.s << "\"ratio\":";
            o.num(els[e]->evaluateRatio());
            o.s << ",";
        }
        o.s << "\"gradient\":";
        o.vec(g);
        o.s << ",\"laplacian\":";
        o.num(els[e]->computeLaplacian());
        o.s << ",\"parameter_gradient\":";
        o.vec(els[e]->computeParameterGradient());
        o.s << "}";
    }
    o.s << "],\"kinetic\":";
    o.num(sys->getKineticEnergy());
    o.s << ",\"external\":";
    o.num(sys->getHamiltonian()->getExternalEnergy());
    o.s << ",\"interaction\":";
    o.num(sys->getInteractionStyle()->getInteractionEnergy());
    o.s << "}";
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s input.in script out.json\n", argv[0]);
        return 2;
    }
    std::string mode = "elements", sampling = "importance";
    int steps = 0, iterations = 0;
    auto* rng = new ScriptedRNG();
    {
        std::ifstream in(argv[2]);
        std::string key;
        while (in >> key) {
            if (key == "mode") in >> mode;
            else if (key == "sampling") in >> sampling;
            else if (key == "steps") in >> steps;
            else if (key == "iterations") in >> iterations;
            else if (key == "positions") {
                std::vector<double> v;
                std::string line;
                std::getline(in, line);
                std::istringstream ls(line);
                double x;
                while (ls >> x) v.push_back(x);
                rng->positions = Eigen::Map<Eigen::VectorXd>(v.data(), (Eigen::Index)v.size());
            } else if (key == "ints") {
                size_t n; in >> n; rng->ints.resize(n);
                for (auto& v : rng->ints) in >> v;
            } else if (key == "gauss") {
                size_t n; in >> n; rng->gauss.resize(n);
                for (auto& v : rng->gauss) in >> v;
            } else if (key == "unif") {
                size_t n; in >> n; rng->unif.resize(n);
                for (auto& v : rng->unif) in >> v;
            } else {
                std::fprintf(stderr, "unknown script key %s\n", key.c_str());
                return 2;
            }
        }
    }

    // same preamble as the reference's main (src/main.cpp:6-12)
    System* sys = new System();
    sys->setNumberOfParticles(2);
    sys->setNumberOfDimensions(2);
    sys->setEquilibrationFraction(0.001);
    char* fake_argv[2] = {argv[0], argv[1]};
    sys->initializeFromConfig(2, fake_argv);
    sys->setRandomNumberGenerator(rng);
    Metropolis* met = nullptr;
    if (sampling == "brute") met = new BruteForce(sys);
    else met = new ImportanceSampling(sys);
    sys->setMetropolis(met);

    // silence the reference's own terminal output while it runs
    std::streambuf* saved = std::cout.rdbuf();
    std::ostringstream sink;
    std::cout.rdbuf(sink.rdbuf());
    sys->initializeSystem();

    Out o;
    o.s << "{\"mode\":\"" << mode << "\",\"snapshots\":[";
    if (mode == "elements") {
        dump_state(o, sys, met, -1, false);
        for (int i = 0; i < steps; i++) {
            sys->runMetropolisCycles();   // numSteps: 1 in the config -> exactly one step
            o.s << ",";
            dump_state(o, sys, met, i, true);
        }
        o.s << "]";
    } else {
        o.s << "]";
        o.s << ",\"iterations\":[";
        Sampler* smp = sys->getSampler();
        Eigen::MatrixXd P = sys->getWeights();
        for (int it = 0; it < iterations; it++) {
            smp->setNumberOfSteps(sys->getStepsWOEqui(), sys->getTotalStepsWOEqui(),
                                  sys->getTotalStepsWEqui());
            sys->runMetropolisCycles();
            smp->computeTotals();
            smp->computeAverages();
            double avgE = smp->getAverageEnergy();
            Eigen::MatrixXd G = smp->getAverageGradients();
            Eigen::MatrixXd GE = smp->getAverageGradientsE();
            Eigen::MatrixXd upd = sys->getOptimization()->updateParameters();
            P -= upd;
            smp->printOutputToTerminal(iterations, 0.0);   // advances the sampler's iteration
            sys->updateAllParameters(P);
            if (it) o.s << ",";
            o.s << "{\"iter\":" << it << ",\"energy\":";
            o.num(avgE);
            o.s << ",\"avg_gradients\":";
            o.mat(G);
            o.s << ",\"avg_gradients_e\":";
            o.mat(GE);
            o.s << ",\"parameters\":";
            o.mat(P);
            o.s << ",\"positions\":";
            o.vec(met->updatePositions());
            o.s << "}";
        }
        o.s << "]";
    }
    o.s << ",\"terminal\":\"";
    for (char ch : sink.str()) {
        if (ch == '\n') o.s << "\\n";
        else if (ch == '"' || ch == '\\') o.s << '\\' << ch;
        else if ((unsigned char)ch >= 0x20 && (unsigned char)ch < 0x80) o.s << ch;
    }
    o.s << "\"}";
    std::cout.rdbuf(saved);
    std::ofstream out(argv[3]);
    out << o.s.str() << "\n";
    MPI_Finalize();
    return 0;
}
