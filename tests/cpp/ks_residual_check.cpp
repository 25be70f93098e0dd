// ks_residual_check.cpp -- the Kohn-Sham residual through the reference-shaped
// C++ API (include/dft_molecule.h): calculateMaximumMolecularKSResidual
// (MolecularSCFHelpers.cpp:119-183) on the GPU.  Reads <dir>/grid.bin (nx ny
// nz xmin xmax ymin ymax zmin zmax), <dir>/potentials.bin (V_alpha then
// V_beta, N values each) and <dir>/orbitals.bin (norb, eigenvalues[norb],
// spins[norb] (0 alpha, 1 beta), then norb columns of N values, all FP64);
// writes ks.out (one value).
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dft_molecule.h"

static std::vector<double> load(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open " + path);
    const std::streamsize bytes = f.tellg();
    f.seekg(0);
    std::vector<double> v((size_t)bytes / sizeof(double));
    if (!f.read(reinterpret_cast<char*>(v.data()), bytes)) throw std::runtime_error("cannot read " + path);
    return v;
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    const std::vector<double> gd = load(dir + "/grid.bin");
    const CartesianGrid grid((std::size_t)gd[0], (std::size_t)gd[1], (std::size_t)gd[2], gd[3], gd[4], gd[5], gd[6],
                             gd[7], gd[8]);
    const std::size_t n = grid.getSize();
    const std::vector<double> pot = load(dir + "/potentials.bin");
    const std::vector<double> raw = load(dir + "/orbitals.bin");
    const std::size_t norb = (std::size_t)raw[0];
    if (pot.size() != 2 * n || raw.size() != 1 + 2 * norb + norb * n) {
        std::fprintf(stderr, "unexpected input sizes\n");
        return 3;
    }
    const std::vector<double> va(pot.begin(), pot.begin() + n), vb(pot.begin() + n, pot.end());
    std::vector<MolecularOrbital> orbs(norb);
    for (std::size_t o = 0; o < norb; ++o) {
        orbs[o].electrons = 1;
        orbs[o].eigenvalue = raw[1 + o];
        orbs[o].spin = raw[1 + norb + o] != 0.0 ? SpinChannel::Beta : SpinChannel::Alpha;
        const double* p = raw.data() + 1 + 2 * norb + o * n;
        orbs[o].psi.assign(p, p + n);
    }
    const double ks = calculateMaximumMolecularKSResidual(grid, va, vb, orbs);
    std::ofstream f(dir + "/ks.out", std::ios::binary);
    f.write(reinterpret_cast<const char*>(&ks), sizeof ks);
    return 0;
}
