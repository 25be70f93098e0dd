// molecule_api.cpp -- the reference's molecular C++ interface (dft_molecule.h)
// served by the GPU kernels.  Each function keeps the reference's argument
// checks, messages and exception types; grid arithmetic goes through the
// C-ABI of libdft_hip.so on a per-thread context.
#include "dft_molecule.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iomanip>
#include <iostream>
#include <memory>
#include <stdexcept>

#include "api_ctx.hpp"
#include "davidson.hpp"
#include "gpu.hpp"
#include "scf.hpp"

namespace {

using dftapi::ctx;
using dftapi::down;
using dftapi::up;

template <class G>
dft_grid to_grid(const G& g) { return dftapi::to_grid(g); }

int device_functional(const XCFunctional& f) {
    if (dynamic_cast<const PZ81*>(&f)) return DFT_XC_PZ81;
    if (dynamic_cast<const PBE96*>(&f)) return DFT_XC_PBE96;
    throw std::invalid_argument("XC functional has no device implementation (PZ81 and PBE96 only)");
}

XCResult eval_point(int id, const XCInput& in) {
    double h_in[4] = {in.alphaDensity, in.betaDensity, in.alphaGradient, in.betaGradient}, h_out[5];
    dftc::Buf di(ctx(), sizeof h_in), dout(ctx(), sizeof h_out);
    di.upload(h_in, sizeof h_in);
    dftc::check(ctx(), dft_xc_points(ctx(), id, di.as<double>(), 1, dout.as<double>()), "dft_xc_points");
    dout.download(h_out, sizeof h_out);
    return {h_out[0], h_out[1], h_out[2], h_out[3], h_out[4]};
}

}  // namespace

// ---- CartesianGrid.cpp:5-159 ------------------------------------------------
CartesianGrid::CartesianGrid(std::size_t nx, std::size_t ny, std::size_t nz, double xmin, double xmax, double ymin,
                             double ymax, double zmin, double zmax)
    : nx_(nx), ny_(ny), nz_(nz), xmin_(xmin), xmax_(xmax), ymin_(ymin), ymax_(ymax), zmin_(zmin), zmax_(zmax) {
    if (nx_ < 2 || ny_ < 2 || nz_ < 2) throw std::invalid_argument("CartesianGrid requires at least two points per dimension.");
    if (xmax_ <= xmin_ || ymax_ <= ymin_ || zmax_ <= zmin_) throw std::invalid_argument("CartesianGrid requires valid spatial bounds.");
    dx_ = (xmax_ - xmin_) / static_cast<double>(nx_ - 1);
    dy_ = (ymax_ - ymin_) / static_cast<double>(ny_ - 1);
    dz_ = (zmax_ - zmin_) / static_cast<double>(nz_ - 1);
}
double CartesianGrid::getX(std::size_t i) const {
    if (i >= nx_) throw std::out_of_range("X grid index out of range.");
    return xmin_ + static_cast<double>(i) * dx_;
}
double CartesianGrid::getY(std::size_t j) const {
    if (j >= ny_) throw std::out_of_range("Y grid index out of range.");
    return ymin_ + static_cast<double>(j) * dy_;
}
double CartesianGrid::getZ(std::size_t k) const {
    if (k >= nz_) throw std::out_of_range("Z grid index out of range.");
    return zmin_ + static_cast<double>(k) * dz_;
}
std::size_t CartesianGrid::getIndex(std::size_t i, std::size_t j, std::size_t k) const {
    if (i >= nx_ || j >= ny_ || k >= nz_) throw std::out_of_range("Cartesian grid index out of range.");
    return (k * ny_ + j) * nx_ + i;
}

// ---- Molecule.cpp:10-56 -----------------------------------------------------
void Molecule::addNucleus(int atomicNumber, double x, double y, double z) {
    if (atomicNumber <= 0) throw std::invalid_argument("Atomic number must be greater than zero.");
    nuclei_.push_back(Nucleus{atomicNumber, {x, y, z}});
}
int Molecule::getElectronCount() const {
    int t = 0;
    for (const Nucleus& n : nuclei_) t += n.atomicNumber;
    return t - charge_;
}

// ---- functionals ----------------------------------------------------------
XCResult PZ81::evaluate(const XCInput& in) const { return eval_point(DFT_XC_PZ81, in); }
XCResult PBE96::evaluate(const XCInput& in) const { return eval_point(DFT_XC_PBE96, in); }

// ---- NuclearRepulsion.cpp:6-54 -------------------------------------------------
double NuclearRepulsion::calculate(const Molecule& molecule) {
    const auto& nu = molecule.getNuclei();
    double e = 0.0;
    for (size_t i = 0; i < nu.size(); ++i)
        for (size_t j = i + 1; j < nu.size(); ++j) {
            const double dx = nu[i].position[0] - nu[j].position[0], dy = nu[i].position[1] - nu[j].position[1],
                         dz = nu[i].position[2] - nu[j].position[2];
            const double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 == 0.0) throw std::runtime_error("NuclearRepulsion is undefined for coincident nuclei.");
            e += (double)nu[i].atomicNumber * (double)nu[j].atomicNumber / std::sqrt(d2);
        }
    return e;
}

// ---- MolecularSCFHelpers.cpp ------------------------------------------------------
double elapsedSeconds(const std::chrono::steady_clock::time_point& a, const std::chrono::steady_clock::time_point& b) {
    return std::chrono::duration<double>(b - a).count();
}

void mixDensity(std::vector<double>& density, const std::vector<double>& output, double mixing) {
    if (density.size() != output.size()) throw std::invalid_argument("Las densidades deben tener el mismo tamano.");
    auto dd = up(density), dout = up(output);
    dftc::check(ctx(), dft_axpby(ctx(), DFT_F64, (int64_t)density.size(), mixing, dout.get(), 1.0 - mixing, dd.get()),
                "dft_axpby");
    density = down(dd, density.size());
}

double calculateMolecularDensityDifference(const CartesianGrid& grid, const std::vector<double>& oldd,
                                           const std::vector<double>& newd) {
    if (oldd.size() != grid.getSize() || newd.size() != grid.getSize())
        throw std::invalid_argument("La malla cartesiana y las densidades deben tener el mismo tamano.");
    auto a = up(oldd), b = up(newd);
    double d2 = 0, n2 = 0;
    dftc::check(ctx(), dft_reduce(ctx(), DFT_RED_DIFF2, DFT_F64, b.get(), a.get(), (int64_t)newd.size(), &d2), "dft_reduce");
    dftc::check(ctx(), dft_reduce(ctx(), DFT_RED_DOT, DFT_F64, b.get(), b.get(), (int64_t)newd.size(), &n2), "dft_reduce");
    const double dv = grid.getDx() * grid.getDy() * grid.getDz();
    const double dn = std::sqrt(d2 * dv), nn = std::sqrt(n2 * dv);
    if (nn <= DFTConstants::EPS) return dn;
    return dn / nn;
}

double calculateMolecularSpinDensityDifference(const CartesianGrid& grid, const std::vector<double>& oa,
                                               const std::vector<double>& ob, const std::vector<double>& na,
                                               const std::vector<double>& nb) {
    return std::max(calculateMolecularDensityDifference(grid, oa, na), calculateMolecularDensityDifference(grid, ob, nb));
}

double calculateMaximumMolecularKSResidual(const CartesianGrid& grid, const std::vector<double>& va,
                                           const std::vector<double>& vb, const std::vector<MolecularOrbital>& orbitals) {
    if (va.size() != grid.getSize() || vb.size() != grid.getSize())
        throw std::invalid_argument("Los potenciales moleculares y la malla deben tener el mismo tamano.");
    const dft_grid g = to_grid(grid);
    const size_t n = grid.getSize();
    auto da = up(va), db = up(vb);
    dftc::Buf p(ctx(), n * 8), hp(ctx(), n * 8), r(ctx(), n * 8);
    dftc::Buf th(ctx(), 8), nr(ctx(), 8);
    double worst = 0.0;
    for (const MolecularOrbital& o : orbitals) {
        if (o.psi.size() != n) throw std::invalid_argument("El orbital molecular y la malla deben tener el mismo tamano.");
        p.upload(o.psi.data(), n * 8);
        th.upload(&o.eigenvalue, 8);
        const void* v = o.spin == SpinChannel::Alpha ? da.get() : db.get();
        dftc::check(ctx(), dft_hpsi(ctx(), &g, DFT_F64, v, p.get(), (int64_t)n, 1, hp.get(), (int64_t)n), "dft_hpsi");
        dftc::check(ctx(), dft_residual(ctx(), DFT_F64, (int64_t)n, 1, hp.get(), (int64_t)n, p.get(), (int64_t)n,
                                        th.as<double>(), r.get(), (int64_t)n), "dft_residual");
        double s = 0;
        dftc::check(ctx(), dft_reduce(ctx(), DFT_RED_DOT, DFT_F64, r.get(), r.get(), (int64_t)n, &s), "dft_reduce");
        worst = std::max(worst, std::sqrt(s * (grid.getDx() * grid.getDy() * grid.getDz())));
    }
    return worst;
}

std::vector<int> buildSpinOccupations(int electronCount) {
    if (electronCount < 0) throw std::invalid_argument("El numero de electrones no puede ser negativo.");
    return std::vector<int>((std::size_t)electronCount, 1);
}

std::vector<MolecularOrbital> convertOrbitalsToSpin(const std::vector<MolecularOrbital>& orbitals, SpinChannel spin) {
    auto c = orbitals;
    for (auto& o : c) o.spin = spin;
    return c;
}

std::vector<MolecularOrbital> solveMolecularOrbitalsSilently(const CartesianGrid& grid, const std::vector<double>& v,
                                                             std::size_t count, const std::vector<int>& occ,
                                                             SpinChannel spin, const std::vector<MolecularOrbital>& init) {
    return solveMolecularOrbitals(grid, v, count, occ, spin, init);  // the block solver writes nothing to stdout
}

void printSCFHeader() {
    std::cout << "\nSCF MOLECULAR\nIter | E (Ha)         | dE         | dRho      | KS        | Mix   | E Alpha      | E Beta\n";
}

void printSCFIteration(int iteration, double energy, double de, double dd, double res, double mix, double,
                       const std::vector<MolecularOrbital>& a, const std::vector<MolecularOrbital>& b, double, double) {
    std::cout << std::setw(4) << iteration << " | " << std::fixed << std::setprecision(10) << energy << " | "
              << std::scientific << std::setprecision(3) << de << " | " << dd << " | " << res << " | " << std::fixed
              << std::setprecision(3) << mix << " | " << std::scientific << std::setprecision(6)
              << (a.empty() ? 0.0 : a.front().eigenvalue) << " | " << (b.empty() ? 0.0 : b.front().eigenvalue) << "\n";
}

void printProfiling(const MolecularSCFTiming& t, double total) {
    std::cout << "initialization " << t.initialization << " s, iterations " << t.totalIterations << " s, mixing "
              << t.mixing << " s, total " << total << " s\n";
}

void printConvergenceAnalysis(const MolecularSCFConvergenceAnalysis& a) {
    std::cout << "last dE/dRho/KS criterion iterations: " << a.lastEnergyCriterionIteration << " "
              << a.lastDensityCriterionIteration << " " << a.lastKSCriterionIteration << "\n";
}

void buildInitialMolecularDensities(const CartesianGrid& grid, const Molecule& molecule, int nal, int nbe, bool closed,
                                    std::vector<double>& ra, std::vector<double>& rb,
                                    std::vector<MolecularOrbital>& ia, std::vector<MolecularOrbital>& ib) {
    const auto vnuc = NuclearPotential::calculate(molecule, grid);
    const auto oa = buildSpinOccupations(nal), ob = buildSpinOccupations(nbe);
    ia = solveMolecularOrbitalsSilently(grid, vnuc, oa.size(), oa, SpinChannel::Alpha, {});
    ib = closed ? convertOrbitalsToSpin(ia, SpinChannel::Beta)
                : solveMolecularOrbitalsSilently(grid, vnuc, ob.size(), ob, SpinChannel::Beta, {});
    ra = calculateMolecularSpinDensity(ia, SpinChannel::Alpha);
    rb = calculateMolecularSpinDensity(ib, SpinChannel::Beta);
}

// ---- MolecularSCFIteration.cpp:12-252 ----------------------------------------------
MolecularSCFIterationResult executeMolecularSCFIteration(const CartesianGrid& grid, const Molecule& molecule,
                                                         const XCFunctional& functional,
                                                         const std::vector<double>& vnuc,
                                                         const std::vector<double>& ra, const std::vector<double>& rb,
                                                         int nal, int nbe, bool closed,
                                                         const std::vector<MolecularOrbital>& pa,
                                                         const std::vector<MolecularOrbital>& pb) {
    const size_t n = grid.getSize();
    if (ra.size() != n || rb.size() != n)
        throw std::invalid_argument("Las densidades moleculares y la malla deben tener el mismo tamano.");
    if (vnuc.size() != n) throw std::invalid_argument("El potencial nuclear y la malla deben tener el mismo tamano.");
    const auto oa = buildSpinOccupations(nal), ob = buildSpinOccupations(nbe);
    MolecularSCFIterationResult it;
    it.density.resize(n);
    for (size_t i = 0; i < n; ++i) it.density[i] = ra[i] + rb[i];
    it.hartreePotential = calculateHartreePotential(grid, it.density);
    it.alphaXCPotential = calculateMolecularSpinExchangeCorrelationPotential(functional, grid, ra, rb, 0);
    it.betaXCPotential = closed ? it.alphaXCPotential
                                : calculateMolecularSpinExchangeCorrelationPotential(functional, grid, ra, rb, 1);
    it.alphaPotential.resize(n);
    it.betaPotential.resize(n);
    for (size_t i = 0; i < n; ++i) {
        it.alphaPotential[i] = vnuc[i] + it.hartreePotential[i] + it.alphaXCPotential[i];
        it.betaPotential[i] = vnuc[i] + it.hartreePotential[i] + it.betaXCPotential[i];
    }
    it.alphaOrbitals = solveMolecularOrbitalsSilently(grid, it.alphaPotential, oa.size(), oa, SpinChannel::Alpha, pa);
    it.betaOrbitals = closed ? convertOrbitalsToSpin(it.alphaOrbitals, SpinChannel::Beta)
                             : solveMolecularOrbitalsSilently(grid, it.betaPotential, ob.size(), ob, SpinChannel::Beta, pb);
    it.orbitals = it.alphaOrbitals;
    it.orbitals.insert(it.orbitals.end(), it.betaOrbitals.begin(), it.betaOrbitals.end());
    it.outputAlphaDensity = calculateMolecularSpinDensity(it.alphaOrbitals, SpinChannel::Alpha);
    it.outputBetaDensity = calculateMolecularSpinDensity(it.betaOrbitals, SpinChannel::Beta);
    it.outputDensity.resize(n);
    for (size_t i = 0; i < n; ++i) it.outputDensity[i] = it.outputAlphaDensity[i] + it.outputBetaDensity[i];
    it.outputHartreePotential = calculateHartreePotential(grid, it.outputDensity);
    it.energy = calculateMolecularTotalEnergy(functional, molecule, grid, it.outputAlphaDensity, it.outputBetaDensity,
                                              it.orbitals, it.outputHartreePotential, vnuc);
    it.densityDifference = calculateMolecularDensityDifference(grid, it.density, it.outputDensity);
    it.spinDensityDifference =
        calculateMolecularSpinDensityDifference(grid, ra, rb, it.outputAlphaDensity, it.outputBetaDensity);
    it.effectiveDensityDifference = std::max(it.densityDifference, it.spinDensityDifference);
    it.residual = calculateMaximumMolecularKSResidual(grid, it.alphaPotential, it.betaPotential, it.orbitals);
    return it;
}

// ---- SelfConsistentFieldMolecule.cpp:50-561 (device-resident loop) -------------------
MolecularResult solveMolecularSelfConsistentField(const CartesianGrid& grid, const Molecule& molecule, int charge,
                                                  int multiplicity, const XCFunctional& functional) {
    if (molecule.getNucleusCount() == 0) throw std::invalid_argument("La molecula debe contener al menos un nucleo.");
    const dft_grid g = to_grid(grid);
    std::vector<double> nuc;
    for (const auto& a : molecule.getNuclei()) nuc.insert(nuc.end(), {(double)a.atomicNumber, a.position[0], a.position[1], a.position[2]});
    dft_scf_config cfg{};
    cfg.functional = device_functional(functional);
    cfg.charge = charge;
    cfg.multiplicity = multiplicity;
    cfg.precision = DFT_SCF_FP64;
    cfg.max_iterations = DFTConstants::MAX_SCF_ITERATIONS;
    cfg.stop_rule = DFT_STOP_REFERENCE;
    cfg.init_guess = DFT_INIT_BARE;
    cfg.davidson_tol = 1.0e-7;
    dftc::ScfDriver drv(ctx(), g, (int)molecule.getNucleusCount(), nuc.data(), cfg);
    dft_scf_report rep;
    dftc::ScfOutputs out;
    out.want_rho = true;
    out.want_fields = true;
    drv.run(rep, out);
    const size_t n = grid.getSize();
    MolecularResult res;
    res.electrons = molecule.getElectronCount() + molecule.getCharge() - charge;
    res.charge = charge;
    SCFResult& s = res.scf;
    const int nal = (int)(out.psi_a.size() / n), nbe = (int)(out.psi_b.size() / n);
    for (int sp = 0; sp < 2; ++sp) {
        const int cnt = sp ? nbe : nal;
        const auto& P = sp ? out.psi_b : out.psi_a;
        for (int i = 0; i < cnt; ++i) {
            MolecularOrbital o;
            o.spin = sp ? SpinChannel::Beta : SpinChannel::Alpha;
            o.electrons = 1;
            o.eigenvalue = out.eig[(size_t)(sp ? nal : 0) + i];
            o.psi.assign(P.begin() + (size_t)i * n, P.begin() + (size_t)(i + 1) * n);
            s.molecularOrbitals.push_back(std::move(o));
        }
    }
    s.alphaDensity.assign(out.rho.begin(), out.rho.begin() + n);
    s.betaDensity.assign(out.rho.begin() + n, out.rho.end());
    s.density.resize(n);
    for (size_t i = 0; i < n; ++i) s.density[i] = s.alphaDensity[i] + s.betaDensity[i];
    s.alphaEffectivePotential = out.veff_a;
    s.betaEffectivePotential = out.veff_b;
    s.effectivePotential = out.veff_a;
    s.energy.kinetic = rep.energy[0];
    s.energy.external = rep.energy[1];
    s.energy.hartree = rep.energy[2];
    s.energy.exchangeCorrelation = rep.energy[3];
    s.energy.nuclearRepulsion = rep.energy[4];
    s.energy.total = rep.energy[5];
    s.densityDifference = rep.density_difference;
    s.energyDifference = rep.energy_difference;
    s.maxKSResidual = rep.ks_residual;
    s.iterations = rep.iterations;
    s.converged = rep.converged != 0;
    return res;
}
