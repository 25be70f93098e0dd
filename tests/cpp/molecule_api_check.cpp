// molecule_api_check.cpp -- exercises the reference-shaped C++ API
// (include/dft_molecule.h, libdft_core.so) the way the reference's SCF driver
// calls it.  Reads raw FP64 inputs from <dir>, writes raw FP64 outputs there;
// tests/test_gpu_cpp_api.py compares them with the oracle.
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dft_molecule.h"

static std::vector<double> load(const std::string& path, std::size_t n) {
    std::vector<double> v(n);
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double))))
        throw std::runtime_error("cannot read " + path);
    return v;
}
static void save(const std::string& path, const std::vector<double>& v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
}

int main(int argc, char** argv) {
    if (argc != 2) return 2;
    const std::string dir = argv[1];
    // grid: nx ny nz xmin xmax ymin ymax zmin zmax
    const std::vector<double> gd = load(dir + "/grid.bin", 9);
    const CartesianGrid grid((std::size_t)gd[0], (std::size_t)gd[1], (std::size_t)gd[2], gd[3], gd[4], gd[5], gd[6],
                             gd[7], gd[8]);
    const std::size_t n = grid.getSize();
    const std::vector<double> v = load(dir + "/v.bin", n), psi = load(dir + "/psi.bin", n);
    const std::vector<double> rho = load(dir + "/rho.bin", n), na = load(dir + "/na.bin", n),
                              nb = load(dir + "/nb.bin", n);
    save(dir + "/hpsi.out", applyMolecularKohnShamHamiltonian(grid, v, psi));
    const std::vector<double> vh = calculateHartreePotential(grid, rho);
    save(dir + "/vh.out", vh);
    const PZ81 pz;
    const PBE96 pbe;
    save(dir + "/vxc_pz.out", calculateMolecularSpinExchangeCorrelationPotential(pz, grid, na, nb, 0));
    save(dir + "/vxc_pbe.out", calculateMolecularSpinExchangeCorrelationPotential(pbe, grid, na, nb, 1));
    Molecule mol;
    mol.addNucleus(8, 0.01, 0.02, 0.03);
    mol.addNucleus(1, 1.43, 1.11, 0.0);
    mol.addNucleus(1, -1.43, 1.11, 0.0);
    save(dir + "/vnuc.out", NuclearPotential::calculate(mol, grid));
    std::vector<double> scal = {calculateHartreeEnergy(grid, rho, vh), calculateMolecularExchangeCorrelationEnergy(pbe, grid, na, nb),
                                NuclearRepulsion::calculate(mol)};
    // Davidson on the nuclear potential: 3 doubly occupied orbitals
    const std::vector<double> vn = NuclearPotential::calculate(mol, grid);
    const std::vector<MolecularOrbital> orbs = solveMolecularOrbitals(grid, vn, 3, {2, 2, 2}, SpinChannel::Alpha);
    for (const auto& o : orbs) scal.push_back(o.eigenvalue);
    save(dir + "/scalars.out", scal);
    save(dir + "/density.out", calculateMolecularSpinDensity(orbs, SpinChannel::Alpha));
    // the reference's own SCF driver entry point: H2 on 7^3 (main.cpp's first case)
    Molecule h2;
    h2.addNucleus(1, -0.7, 0.0, 0.0);
    h2.addNucleus(1, 0.7, 0.0, 0.0);
    const CartesianGrid g7(7, 7, 7, -8.1, 7.9, -8.1, 7.9, -8.1, 7.9);
    const MolecularResult r = solveMolecularSelfConsistentField(g7, h2, 0, 1, pz);
    save(dir + "/scf.out", {r.scf.energy.total, (double)r.scf.iterations, r.scf.converged ? 1.0 : 0.0});
    std::printf("molecule_api_check: ok\n");
    return 0;
}
