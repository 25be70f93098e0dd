// ref_shim.cpp -- extern "C" wrappers around the compiled reference
// (willandru/DFT src/dft/*.cpp, built from where they lie by oracle/Makefile
// into oracle/_ref/).  TEST INFRASTRUCTURE ONLY: used to write the golden
// fixtures under tests/golden and to time the reference CPU path.  Each
// wrapper calls the reference's public function named in its comment and
// turns exceptions into status codes (-1 invalid_argument, -2 other).
#include "AtomicDFT.h"
#include "CartesianGrid.h"
#include "DFTData.h"
#include "ElectronicConfiguration.h"
#include "EigenvalueSolverMolecule.h"
#include "ElectronDensityMolecule.h"
#include "ExchangeCorrelationMolecule.h"
#include "HartreePotentialMolecule.h"
#include "KohnShamHamiltonianMolecule.h"
#include "MolecularSCFHelpers.h"
#include "MolecularSCFIteration.h"
#include "Molecule.h"
#include "NuclearPotential.h"
#include "NuclearRepulsion.h"
#include "PBE96.h"
#include "PZ81.h"
#include "RadialGrid.h"
#include "SelfConsistentFieldMolecule.h"
#include "TotalEnergyMolecule.h"

#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {
std::string g_err;

CartesianGrid make_grid(const long* n, const double* b) {
    return CartesianGrid(n[0], n[1], n[2], b[0], b[1], b[2], b[3], b[4], b[5]);
}
std::vector<double> vec(const double* p, size_t n) { return std::vector<double>(p, p + n); }

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return -1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -2;
    }
}

const XCFunctional& functional(int id) {
    static PZ81 pz;
    static PBE96 pbe;
    if (id == 1) return pz;
    if (id == 2) return pbe;
    throw std::invalid_argument("unknown functional id");
}

Molecule make_molecule(int natoms, const double* nuc, int charge) {
    Molecule m(charge);
    for (int a = 0; a < natoms; ++a) m.addNucleus((int)nuc[4 * a], nuc[4 * a + 1], nuc[4 * a + 2], nuc[4 * a + 3]);
    return m;
}
}  // namespace

extern "C" {

const char* rf_last_error() { return g_err.c_str(); }

// grid n = {nx, ny, nz}; bounds b = {xmin, xmax, ymin, ymax, zmin, zmax}
int rf_grid_spacing(const long* n, const double* b, double* d) {
    return guarded([&] {
        auto g = make_grid(n, b);
        d[0] = g.getDx(); d[1] = g.getDy(); d[2] = g.getDz();
    });
}

// applyMolecularKohnShamHamiltonian
int rf_hpsi(const long* n, const double* b, const double* v, const double* psi, double* out) {
    return guarded([&] {
        auto g = make_grid(n, b);
        auto r = applyMolecularKohnShamHamiltonian(g, vec(v, g.getSize()), vec(psi, g.getSize()));
        std::memcpy(out, r.data(), r.size() * sizeof(double));
    });
}

// calculateHartreePotential / calculateHartreeEnergy
int rf_poisson(const long* n, const double* b, const double* rho, double* out) {
    return guarded([&] {
        auto g = make_grid(n, b);
        auto r = calculateHartreePotential(g, vec(rho, g.getSize()));
        std::memcpy(out, r.data(), r.size() * sizeof(double));
    });
}
int rf_hartree_energy(const long* n, const double* b, const double* rho, const double* vh, double* e) {
    return guarded([&] {
        auto g = make_grid(n, b);
        *e = calculateHartreeEnergy(g, vec(rho, g.getSize()), vec(vh, g.getSize()));
    });
}

// XCFunctional::evaluate (PZ81 / PBE96) over npts points of {na, nb, ga, gb}
int rf_xc_points(int id, long npts, const double* in, double* out) {
    return guarded([&] {
        const XCFunctional& f = functional(id);
        for (long p = 0; p < npts; ++p) {
            XCInput x{in[4 * p], in[4 * p + 1], in[4 * p + 2], in[4 * p + 3]};
            XCResult r = f.evaluate(x);
            double* o = out + 5 * p;
            o[0] = r.energyPerElectron; o[1] = r.potentialAlpha; o[2] = r.potentialBeta;
            o[3] = r.gradientCoefficientAlpha; o[4] = r.gradientCoefficientBeta;
        }
    });
}

// calculateMolecularSpinExchangeCorrelationPotential / ...Energy
int rf_xc_potential(const long* n, const double* b, int id, const double* na, const double* nb, int spin, double* out) {
    return guarded([&] {
        auto g = make_grid(n, b);
        auto r = calculateMolecularSpinExchangeCorrelationPotential(functional(id), g, vec(na, g.getSize()),
                                                                    vec(nb, g.getSize()), spin);
        std::memcpy(out, r.data(), r.size() * sizeof(double));
    });
}
int rf_xc_energy(const long* n, const double* b, int id, const double* na, const double* nb, double* e) {
    return guarded([&] {
        auto g = make_grid(n, b);
        *e = calculateMolecularExchangeCorrelationEnergy(functional(id), g, vec(na, g.getSize()), vec(nb, g.getSize()));
    });
}

// calculateMolecularSpinDensity (alpha orbitals, given electrons)
int rf_spin_density(long npts, int norb, const double* psi, const int* electrons, double* out) {
    return guarded([&] {
        std::vector<MolecularOrbital> orbs(norb);
        for (int o = 0; o < norb; ++o) {
            orbs[o].spin = SpinChannel::Alpha;
            orbs[o].electrons = electrons[o];
            orbs[o].psi = vec(psi + (size_t)o * npts, npts);
        }
        auto r = calculateMolecularSpinDensity(orbs, SpinChannel::Alpha);
        std::memcpy(out, r.data(), r.size() * sizeof(double));
    });
}

// calculateMolecularElectronDensity (spins[o]: 0 alpha, 1 beta)
int rf_electron_density(long npts, int norb, const double* psi, const int* electrons, const int* spins, double* out) {
    return guarded([&] {
        std::vector<MolecularOrbital> orbs(norb);
        for (int o = 0; o < norb; ++o) {
            orbs[o].spin = spins[o] ? SpinChannel::Beta : SpinChannel::Alpha;
            orbs[o].electrons = electrons[o];
            orbs[o].psi = vec(psi + (size_t)o * npts, npts);
        }
        auto r = calculateMolecularElectronDensity(orbs);
        std::memcpy(out, r.data(), r.size() * sizeof(double));
    });
}

// integrateMolecularElectronDensity
int rf_integrate_density(long npts, const double* rho, double dx, double dy, double dz, double* e) {
    return guarded([&] { *e = integrateMolecularElectronDensity(vec(rho, npts), dx, dy, dz); });
}

// calculateMolecularKineticEnergy / calculateMolecularExternalEnergy
int rf_kinetic(const long* n, const double* b, int norb, const double* psi, const int* electrons, double* e) {
    return guarded([&] {
        auto g = make_grid(n, b);
        std::vector<MolecularOrbital> orbs(norb);
        for (int o = 0; o < norb; ++o) {
            orbs[o].electrons = electrons[o];
            orbs[o].psi = vec(psi + (size_t)o * g.getSize(), g.getSize());
        }
        *e = calculateMolecularKineticEnergy(g, orbs);
    });
}
int rf_external(const long* n, const double* b, const double* rho, const double* vnuc, double* e) {
    return guarded([&] {
        auto g = make_grid(n, b);
        *e = calculateMolecularExternalEnergy(g, vec(rho, g.getSize()), vec(vnuc, g.getSize()));
    });
}

// NuclearPotential::calculate / NuclearRepulsion::calculate
int rf_nuclear_potential(const long* n, const double* b, int natoms, const double* nuc, double* out) {
    return guarded([&] {
        auto g = make_grid(n, b);
        auto r = NuclearPotential::calculate(make_molecule(natoms, nuc, 0), g);
        std::memcpy(out, r.data(), r.size() * sizeof(double));
    });
}
int rf_nuclear_repulsion(int natoms, const double* nuc, double* e) {
    return guarded([&] { *e = NuclearRepulsion::calculate(make_molecule(natoms, nuc, 0)); });
}

// solveMolecularOrbitals (alpha, occupations 1); init may be null.
int rf_davidson(const long* n, const double* b, const double* v, int norb, const double* init, long maxit,
                double* psi_out, double* eig_out) {
    return guarded([&] {
        auto g = make_grid(n, b);
        std::vector<MolecularOrbital> initial;
        if (init) {
            initial.resize(norb);
            for (int o = 0; o < norb; ++o) initial[o].psi = vec(init + (size_t)o * g.getSize(), g.getSize());
        }
        std::vector<int> occ(norb, 1);
        std::ostringstream sink;
        std::streambuf* old = std::cerr.rdbuf(sink.rdbuf());
        try {
            auto r = solveMolecularOrbitals(g, vec(v, g.getSize()), norb, occ, SpinChannel::Alpha, initial, maxit);
            std::cerr.rdbuf(old);
            for (int o = 0; o < norb; ++o) {
                std::memcpy(psi_out + (size_t)o * g.getSize(), r[o].psi.data(), g.getSize() * sizeof(double));
                eig_out[o] = r[o].eigenvalue;
            }
        } catch (...) {
            std::cerr.rdbuf(old);
            throw;
        }
    });
}

// executeMolecularSCFIteration on given spin densities, orbitals started from
// the seed vectors (no previous orbitals).  Outputs as og_scf_step: vh, vxa,
// vxb, ora, orb, ovh (N each), eig (alpha then beta), E[6], dd (effective
// density difference), ks (max KS residual).
// calculateMaximumMolecularKSResidual (MolecularSCFHelpers.cpp:119-183); spins[o]: 0 alpha, 1 beta
int rf_ks_residual(const long* n, const double* b, const double* va, const double* vb, int norb, const double* psi,
                   const double* eig, const int* spins, double* out) {
    return guarded([&] {
        auto g = make_grid(n, b);
        const size_t N = g.getSize();
        std::vector<MolecularOrbital> orbs((size_t)norb);
        for (int o = 0; o < norb; ++o) {
            orbs[o].spin = spins[o] ? SpinChannel::Beta : SpinChannel::Alpha;
            orbs[o].electrons = 1;
            orbs[o].eigenvalue = eig[o];
            orbs[o].psi = vec(psi + (size_t)o * N, N);
        }
        *out = calculateMaximumMolecularKSResidual(g, vec(va, N), vec(vb, N), orbs);
    });
}

int rf_scf_step(const long* n, const double* b, int natoms, const double* nuc, int id, int mult, const double* ra,
                const double* rb, double* vh, double* vxa, double* vxb, double* ora, double* orb, double* ovh,
                double* eig, double* E, double* dd, double* ks) {
    return guarded([&] {
        auto g = make_grid(n, b);
        Molecule m = make_molecule(natoms, nuc, 0);
        const int ne = m.getElectronCount(), sd = mult - 1;
        const int na = (ne + sd) / 2, nbeta = (ne - sd) / 2;
        const size_t N = g.getSize();
        const std::vector<double> vn = NuclearPotential::calculate(m, g);
        std::ostringstream sink1, sink2;
        std::streambuf* o1 = std::cout.rdbuf(sink1.rdbuf());
        std::streambuf* o2 = std::cerr.rdbuf(sink2.rdbuf());
        try {
            const MolecularSCFIterationResult r = executeMolecularSCFIteration(
                g, m, functional(id), vn, vec(ra, N), vec(rb, N), na, nbeta, mult == 1, {}, {});
            std::cout.rdbuf(o1);
            std::cerr.rdbuf(o2);
            const std::vector<double>* fields[6] = {&r.hartreePotential, &r.alphaXCPotential, &r.betaXCPotential,
                                                    &r.outputAlphaDensity, &r.outputBetaDensity, &r.outputHartreePotential};
            double* outs[6] = {vh, vxa, vxb, ora, orb, ovh};
            for (int f = 0; f < 6; ++f) std::memcpy(outs[f], fields[f]->data(), N * sizeof(double));
            for (int o = 0; o < na; ++o) eig[o] = r.alphaOrbitals[o].eigenvalue;
            for (int o = 0; o < nbeta; ++o) eig[na + o] = r.betaOrbitals[o].eigenvalue;
            const double e[6] = {r.energy.kinetic, r.energy.external, r.energy.hartree, r.energy.exchangeCorrelation,
                                 r.energy.nuclearRepulsion, r.energy.total};
            std::memcpy(E, e, sizeof e);
            *dd = r.effectiveDensityDifference;
            *ks = r.residual;
        } catch (...) {
            std::cout.rdbuf(o1);
            std::cerr.rdbuf(o2);
            throw;
        }
    });
}

// solveMolecularSelfConsistentField.  out = {iterations, converged, T, Eext,
// EH, Exc, Enn, Etot, dE, dRho, KS}; eig_out: all orbital eigenvalues (alpha
// then beta), capacity neig.
int rf_scf(const long* n, const double* b, int natoms, const double* nuc, int charge, int mult, int id,
           double* out, double* eig_out, int neig, int quiet) {
    return guarded([&] {
        auto g = make_grid(n, b);
        Molecule m = make_molecule(natoms, nuc, charge);
        std::ostringstream sink1, sink2;
        std::streambuf* o1 = nullptr;
        std::streambuf* o2 = nullptr;
        if (quiet) {
            o1 = std::cout.rdbuf(sink1.rdbuf());
            o2 = std::cerr.rdbuf(sink2.rdbuf());
        }
        try {
            MolecularResult r = solveMolecularSelfConsistentField(g, m, charge, mult, functional(id));
            if (quiet) { std::cout.rdbuf(o1); std::cerr.rdbuf(o2); }
            const auto& s = r.scf;
            double v[11] = {(double)s.iterations, s.converged ? 1.0 : 0.0, s.energy.kinetic, s.energy.external,
                            s.energy.hartree, s.energy.exchangeCorrelation, s.energy.nuclearRepulsion,
                            s.energy.total, s.energyDifference, s.densityDifference, s.maxKSResidual};
            std::memcpy(out, v, sizeof v);
            for (int i = 0; i < neig && i < (int)s.molecularOrbitals.size(); ++i) eig_out[i] = s.molecularOrbitals[i].eigenvalue;
        } catch (...) {
            if (quiet) { std::cout.rdbuf(o1); std::cerr.rdbuf(o2); }
            throw;
        }
    });
}

// solveAtom (AtomicDFT.cpp:6-41) on RadialGrid(points, rmax) with
// getAtomicConfiguration(Z): out = {iterations, converged, kinetic, external,
// hartree, xc, total, dE, dRho, KS residual}; eig: the orbitals' eigenvalues in
// the result's order (up to neig)
int rf_radial_scf(int Z, int fid, long points, double rmax, double* out, double* eig, int neig, int* norb) {
    return guarded([&] {
        const RadialGrid grid((std::size_t)points, rmax);
        const AtomicResult r = solveAtom(grid, getAtomicConfiguration(Z), functional(fid));
        const auto& s = r.scf;
        double v[10] = {(double)s.iterations, s.converged ? 1.0 : 0.0, s.energy.kinetic, s.energy.external,
                        s.energy.hartree, s.energy.exchangeCorrelation, s.energy.total, s.energyDifference,
                        s.densityDifference, s.maxKSResidual};
        std::memcpy(out, v, sizeof v);
        *norb = (int)s.orbitals.size();
        for (int i = 0; i < neig && i < (int)s.orbitals.size(); ++i) eig[i] = s.orbitals[i].eigenvalue;
    });
}

}  // extern "C"
