// density_api_check.cpp -- the total electron density and its integral
// through the reference-shaped C++ API (include/dft_molecule.h):
// calculateMolecularElectronDensity and integrateMolecularElectronDensity,
// with their argument errors.  Reads <dir>/grid.bin (nx ny nz xmin xmax ymin
// ymax zmin zmax), <dir>/orbitals.bin (norb, electrons[norb], spins[norb],
// then norb columns of N values, all FP64); writes density.out,
// integral.out and errors.out (1 for each expected exception seen).
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
static void save(const std::string& path, const std::vector<double>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}

template <class E, class F>
static double raises(F&& f) {
    try {
        f();
    } catch (const E&) {
        return 1.0;
    } catch (...) {
        return 0.0;
    }
    return 0.0;
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    const std::vector<double> gd = load(dir + "/grid.bin");
    const CartesianGrid grid((std::size_t)gd[0], (std::size_t)gd[1], (std::size_t)gd[2], gd[3], gd[4], gd[5], gd[6],
                             gd[7], gd[8]);
    const std::size_t n = grid.getSize();
    const std::vector<double> raw = load(dir + "/orbitals.bin");
    const std::size_t norb = (std::size_t)raw[0];
    if (raw.size() != 1 + 2 * norb + norb * n) {
        std::fprintf(stderr, "orbitals.bin holds %zu values, expected %zu\n", raw.size(), 1 + 2 * norb + norb * n);
        return 3;
    }
    std::vector<MolecularOrbital> orbs(norb);
    for (std::size_t o = 0; o < norb; ++o) {
        orbs[o].electrons = (int)raw[1 + o];
        orbs[o].spin = raw[1 + norb + o] != 0.0 ? SpinChannel::Beta : SpinChannel::Alpha;
        const double* p = raw.data() + 1 + 2 * norb + o * n;
        orbs[o].psi.assign(p, p + n);
    }
    const std::vector<double> density = calculateMolecularElectronDensity(orbs);
    save(dir + "/density.out", density);
    save(dir + "/integral.out", {integrateMolecularElectronDensity(density, grid.getDx(), grid.getDy(), grid.getDz())});

    // argument errors (ElectronDensityMolecule.cpp:9-13, 93-97, 122-132)
    std::vector<MolecularOrbital> ragged = orbs;
    for (auto& o : ragged)
        if (o.spin == SpinChannel::Beta) o.psi.resize(n - 1);
    std::vector<double> errs = {
        raises<std::invalid_argument>([&] { integrateMolecularElectronDensity({}, 0.1, 0.1, 0.1); }),
        raises<std::invalid_argument>([&] { integrateMolecularElectronDensity(density, 0.1, 0.0, 0.1); }),
        raises<std::invalid_argument>([&] { calculateMolecularElectronDensity({}); }),
        raises<std::runtime_error>([&] { calculateMolecularElectronDensity(ragged); }),
    };
    save(dir + "/errors.out", errs);
    return 0;
}
