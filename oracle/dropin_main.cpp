// dropin_main.cpp -- checker program: the reference's own molecular SCF driver
// (SelfConsistentFieldMolecule.cpp, MolecularSCFIteration.cpp,
// MolecularSCFHelpers.cpp, compiled unchanged from the reference's tree)
// running over the GPU hot path (dft_amd/host/ks_path.cpp built with
// -DDFT_REFERENCE_HEADERS in place of the reference's hot-path files).
// Written against the reference's headers, like main.cpp:69-331.  Prints one
// line per system: name, total energy, SCF iterations, converged flag.
#include "CartesianGrid.h"
#include "Molecule.h"
#include "PBE96.h"
#include "PZ81.h"
#include "SelfConsistentFieldMolecule.h"

#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>

int main(int argc, char** argv) {
    const char* which = argc > 1 ? argv[1] : "h2";
    PZ81 pz81;
    PBE96 pbe96;
    // the driver prints its iteration table on stdout; keep stdout for the result lines
    std::ostringstream sink;
    std::streambuf* keep = std::cout.rdbuf(sink.rdbuf());
    MolecularResult r;
    try {
        if (!std::strcmp(which, "h2")) {  // main.cpp:74-117
            Molecule m(0);
            m.addNucleus(1, -0.7, 0.0, 0.0);
            m.addNucleus(1, 0.7, 0.0, 0.0);
            CartesianGrid g(7, 7, 7, -8.1, 7.9, -8.1, 7.9, -8.1, 7.9);
            r = solveMolecularSelfConsistentField(g, m, m.getCharge(), 1, pz81);
        } else if (!std::strcmp(which, "h2o")) {  // main.cpp:119-175
            Molecule m(0);
            m.addNucleus(8, 0.0, 0.0, 0.0);
            m.addNucleus(1, 1.430, 0.0, 1.108);
            m.addNucleus(1, -1.430, 0.0, 1.108);
            CartesianGrid g(9, 9, 9, -8.1, 7.9, -8.1, 7.9, -8.1, 7.9);
            r = solveMolecularSelfConsistentField(g, m, m.getCharge(), 1, pz81);
        } else if (!std::strcmp(which, "he_pbe")) {  // He, PBE96, 16^3, h = 0.5 (the PBE dispatch)
            Molecule m(0);
            m.addNucleus(2, 0.0, 0.0, 0.0);
            CartesianGrid g(16, 16, 16, -3.75, 3.75, -3.75, 3.75, -3.75, 3.75);
            r = solveMolecularSelfConsistentField(g, m, m.getCharge(), 1, pbe96);
        } else {
            std::cout.rdbuf(keep);
            std::fprintf(stderr, "unknown system %s\n", which);
            return 2;
        }
    } catch (const std::exception& e) {
        std::cout.rdbuf(keep);
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    std::cout.rdbuf(keep);
    std::printf("%s %.15e %d %d\n", which, r.scf.energy.total, r.scf.iterations, r.scf.converged ? 1 : 0);
    return 0;
}
