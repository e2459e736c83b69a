"""Applies the engine binding (what a maintainer adds to the reference so that
its System drives the engine of include/vmc.h) to a copy of the reference's
sources: copies REF/src to OUT/src and edits the copy (test
infrastructure; OUT lies under oracle/_ref, which is never committed).

Edits, each located by a declaration or call the reference's files hold:
  src/system.h   : #include "vmc.h"; four public members of System
  src/sampler.h  : #include "vmc.h"; Sampler::setTotals next to computeTotals
  src/system.cpp : runSimulation calls initializeEngine after initializeSystem,
                   runEngineCycles instead of runMetropolisCycles and
                   computeTotals, and passes the updated parameters on;
                   system_engine.inc appended
  src/sampler.cpp: sampler_totals.inc appended

    python3 apply.py REF OUT
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def edit(path, old, new, count=1):
    text = open(path).read()
    n = text.count(old)
    if n != count:
        raise SystemExit(f"{path}: expected {count} occurrence(s) of {old!r}, found {n}")
    open(path, "w").write(text.replace(old, new))


def main(ref, out):
    src = os.path.join(out, "src")
    if os.path.isdir(src):
        shutil.rmtree(src)
    shutil.copytree(os.path.join(ref, "src"), src)
    for d, _, files in os.walk(src):   # the copy is ours to edit
        os.chmod(d, 0o755)
        for f in files:
            os.chmod(os.path.join(d, f), 0o644)
    # system.h: the engine handle and the three binding calls
    edit(os.path.join(src, "system.h"), '#include "main.h"',
         '#include "main.h"\n#include "vmc.h"')
    edit(os.path.join(src, "system.h"), "public:",
         "public:\n    void initializeEngine();\n    void engineSetParameters();\n"
         "    void runEngineCycles();\n    vmc_engine* m_engine = nullptr;\n"
         "    int m_engineWalkers = 1;")
    # sampler.h: setTotals beside computeTotals
    edit(os.path.join(src, "sampler.h"), "#include <Eigen/Dense>",
         '#include <Eigen/Dense>\n#include <vector>\n#include "vmc.h"')
    edit(os.path.join(src, "sampler.h"), "    void computeTotals();",
         "    void computeTotals();\n    void setTotals(const vmc_sums& s, const std::vector<double>& g,"
         "\n                   const std::vector<double>& ge);")
    # system.cpp: the iteration loop drives the engine
    cpp = os.path.join(src, "system.cpp")
    edit(cpp, "    initializeSystem();\n", "    initializeSystem();\n    initializeEngine();\n")
    edit(cpp, "        runMetropolisCycles();\n", "        runEngineCycles();\n")
    edit(cpp, "        m_sampler->computeTotals();\n", "")
    edit(cpp, "        updateAllParameters(m_parameters);\n",
         "        updateAllParameters(m_parameters);\n        engineSetParameters();\n")
    with open(cpp, "a") as f:
        f.write("\n" + open(os.path.join(HERE, "system_engine.inc")).read())
    with open(os.path.join(src, "sampler.cpp"), "a") as f:
        f.write("\n" + open(os.path.join(HERE, "sampler_totals.inc")).read())


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
