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

#include "davidson.hpp"
#include "gpu.hpp"
#include "scf.hpp"

namespace {

thread_local int t_device = 0;
thread_local std::unique_ptr<dftc::Ctx> t_ctx;

dft_ctx* ctx() {
    if (!t_ctx) t_ctx.reset(new dftc::Ctx(t_device));
    return t_ctx->get();
}

dft_grid to_grid(const CartesianGrid& g) {
    dft_grid d;
    if (dft_grid_init(&d, (int64_t)g.getNx(), (int64_t)g.getNy(), (int64_t)g.getNz(), g.getXMin(), g.getXMax(),
                      g.getYMin(), g.getYMax(), g.getZMin(), g.getZMax()) != DFT_OK)
        throw std::invalid_argument("CartesianGrid requires valid spatial bounds.");
    return d;
}

// device copy of a host vector
dftc::Buf up(const std::vector<double>& v) {
    dftc::Buf b(ctx(), std::max<size_t>(v.size(), 1) * sizeof(double));
    if (!v.empty()) b.upload(v.data(), v.size() * sizeof(double));
    return b;
}

std::vector<double> down(const dftc::Buf& b, size_t n) {
    std::vector<double> v(n);
    if (n) b.download(v.data(), n * sizeof(double));
    return v;
}

int device_functional(const XCFunctional& f) {
    const int id = f.deviceFunctional();
    if (id != DFT_XC_PZ81 && id != DFT_XC_PBE96)
        throw std::invalid_argument("XC functional has no device implementation (PZ81 and PBE96 only)");
    return id;
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

namespace dft {
void setDevice(int device) {
    if (t_device != device) t_ctx.reset();
    t_device = device;
}
}  // namespace dft

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

// ---- KohnShamHamiltonianMolecule.cpp:7-183 -----------------------------------
std::vector<double> applyMolecularKohnShamHamiltonian(const CartesianGrid& grid, const std::vector<double>& veff,
                                                      const std::vector<double>& psi) {
    if (veff.size() != grid.getSize())
        throw std::invalid_argument("El potencial efectivo y la malla cartesiana deben tener el mismo tamano.");
    if (psi.size() != grid.getSize())
        throw std::invalid_argument("La funcion de onda y la malla cartesiana deben tener el mismo tamano.");
    const dft_grid g = to_grid(grid);
    auto dv = up(veff), dp = up(psi);
    dftc::Buf out(ctx(), psi.size() * sizeof(double));
    dftc::check(ctx(), dft_hpsi(ctx(), &g, DFT_F64, dv.get(), dp.get(), (int64_t)psi.size(), 1, out.get(),
                                (int64_t)psi.size()), "dft_hpsi");
    return down(out, psi.size());
}

// ---- HartreePotentialMolecule.cpp:99-379 ---------------------------------------
std::vector<double> calculateHartreePotential(const CartesianGrid& grid, const std::vector<double>& density) {
    if (density.size() != grid.getSize())
        throw std::invalid_argument("La densidad molecular y la malla cartesiana deben tener el mismo tamano.");
    if (density.empty()) throw std::invalid_argument("La densidad molecular no puede estar vacia.");
    const dft_grid g = to_grid(grid);
    auto dr = up(density);
    dftc::Buf v(ctx(), density.size() * sizeof(double));
    int64_t sweeps = 0;
    double md = 0;
    dftc::check(ctx(), dft_poisson(ctx(), &g, DFT_F64, dr.get(), v.get(), 10000, 1.0e-8, 1, 64, &sweeps, &md),
                "dft_poisson");
    return down(v, density.size());
}

double calculateHartreeEnergy(const CartesianGrid& grid, const std::vector<double>& density,
                              const std::vector<double>& vh) {
    if (density.size() != grid.getSize() || vh.size() != grid.getSize())
        throw std::invalid_argument("La densidad, el potencial de Hartree y la malla deben tener el mismo tamano.");
    if (density.empty()) throw std::invalid_argument("La densidad molecular no puede estar vacia.");
    const dft_grid g = to_grid(grid);
    auto dr = up(density), dvh = up(vh);
    double e = 0;
    dftc::check(ctx(), dft_hartree_energy(ctx(), &g, DFT_F64, dr.get(), dvh.get(), &e), "dft_hartree_energy");
    return e;
}

// ---- ExchangeCorrelationMolecule.cpp:576-842 -----------------------------------
std::vector<double> calculateMolecularSpinExchangeCorrelationPotential(const XCFunctional& functional,
                                                                       const CartesianGrid& grid,
                                                                       const std::vector<double>& na,
                                                                       const std::vector<double>& nb, int spin) {
    if (spin != 0 && spin != 1) throw std::invalid_argument("El canal de spin debe ser 0 (alpha) o 1 (beta).");
    if (na.size() != grid.getSize() || nb.size() != grid.getSize())
        throw std::invalid_argument("La malla cartesiana y las densidades spin deben tener el mismo tamano.");
    const int id = device_functional(functional);
    const dft_grid g = to_grid(grid);
    auto da = up(na), db = up(nb);
    dftc::Buf va(ctx(), na.size() * 8), vb(ctx(), na.size() * 8);
    dftc::check(ctx(), dft_xc(ctx(), &g, DFT_F64, id, da.get(), db.get(), va.get(), spin == 1 ? vb.get() : nullptr, nullptr),
                "dft_xc");
    return down(spin == 0 ? va : vb, na.size());
}

double calculateMolecularExchangeCorrelationEnergy(const XCFunctional& functional, const CartesianGrid& grid,
                                                   const std::vector<double>& na, const std::vector<double>& nb) {
    if (na.size() != grid.getSize() || nb.size() != grid.getSize())
        throw std::invalid_argument("La malla cartesiana y las densidades spin deben tener el mismo tamano.");
    const int id = device_functional(functional);
    const dft_grid g = to_grid(grid);
    auto da = up(na), db = up(nb);
    dftc::Buf va(ctx(), na.size() * 8);
    double e = 0;
    dftc::check(ctx(), dft_xc(ctx(), &g, DFT_F64, id, da.get(), db.get(), va.get(), nullptr, &e), "dft_xc");
    return e;
}

// ---- ElectronDensityMolecule.cpp:5-76 ------------------------------------------
namespace {
// calculateMolecularSpinDensity's checks and sum (ElectronDensityMolecule.cpp:5-76), leaving
// the density on the device; n receives its length.
dftc::Buf spin_density_device(const std::vector<MolecularOrbital>& orbitals, SpinChannel spin, size_t& n) {
    if (orbitals.empty()) throw std::invalid_argument("La lista de orbitales moleculares no puede estar vacia.");
    n = 0;
    bool init = false;
    std::vector<const MolecularOrbital*> occ;
    for (const MolecularOrbital& o : orbitals) {
        if (o.spin != spin || o.electrons <= 0) continue;
        if (o.psi.empty())
            throw std::invalid_argument("Un orbital molecular ocupado no puede tener una funcion de onda vacia.");
        if (!init) { n = o.psi.size(); init = true; }
        else if (o.psi.size() != n)
            throw std::invalid_argument("Todos los orbitales moleculares deben estar definidos sobre la misma malla.");
        occ.push_back(&o);
    }
    if (!init)
        for (const MolecularOrbital& o : orbitals)
            if (!o.psi.empty()) { n = o.psi.size(); init = true; break; }
    if (!init) throw std::invalid_argument("No existen funciones de onda moleculares validas.");
    dftc::Buf out(ctx(), n * 8);
    if (occ.empty()) {
        out.zero();
        return out;
    }
    std::vector<double> block(occ.size() * n), el(occ.size());
    for (size_t i = 0; i < occ.size(); ++i) {
        std::copy(occ[i]->psi.begin(), occ[i]->psi.end(), block.begin() + i * n);
        el[i] = (double)occ[i]->electrons;
    }
    auto dp = up(block);
    dftc::check(ctx(), dft_density(ctx(), DFT_F64, dp.get(), (int64_t)n, (int)occ.size(), el.data(), (int64_t)n, out.get(), 0),
                "dft_density");
    return out;
}
}  // namespace

std::vector<double> calculateMolecularSpinDensity(const std::vector<MolecularOrbital>& orbitals, SpinChannel spin) {
    size_t n = 0;
    dftc::Buf d = spin_density_device(orbitals, spin, n);
    return down(d, n);
}

std::vector<double> calculateMolecularElectronDensity(const std::vector<MolecularOrbital>& orbitals) {
    size_t na = 0, nb = 0;
    dftc::Buf da = spin_density_device(orbitals, SpinChannel::Alpha, na);
    dftc::Buf db = spin_density_device(orbitals, SpinChannel::Beta, nb);
    if (na != nb) throw std::runtime_error("Las densidades alpha y beta deben tener el mismo tamano.");
    // alpha + beta pointwise: 1 * beta + 1 * alpha rounds exactly as alpha + beta
    dftc::check(ctx(), dft_axpby(ctx(), DFT_F64, (int64_t)na, 1.0, db.get(), 1.0, da.get()), "dft_axpby");
    return down(da, na);
}

double integrateMolecularElectronDensity(const std::vector<double>& density, double dx, double dy, double dz) {
    if (density.empty()) throw std::invalid_argument("La densidad molecular no puede estar vacia.");
    if (dx <= 0.0 || dy <= 0.0 || dz <= 0.0)
        throw std::invalid_argument("Los pasos de la malla cartesiana deben ser mayores que cero.");
    auto d = up(density);
    double s = 0.0;
    dftc::check(ctx(), dft_reduce(ctx(), DFT_RED_SUM, DFT_F64, d.get(), nullptr, (int64_t)density.size(), &s), "dft_reduce");
    return s * dx * dy * dz;
}

// ---- TotalEnergyMolecule.cpp:13-333 --------------------------------------------
double calculateMolecularKineticEnergy(const CartesianGrid& grid, const std::vector<MolecularOrbital>& orbitals) {
    const dft_grid g = to_grid(grid);
    const size_t n = grid.getSize();
    std::vector<double> block;
    std::vector<int> el;
    for (const MolecularOrbital& o : orbitals) {
        if (o.electrons <= 0) continue;
        if (o.psi.size() != n) throw std::invalid_argument("El orbital molecular y la malla deben tener el mismo tamano.");
        block.insert(block.end(), o.psi.begin(), o.psi.end());
        el.push_back(o.electrons);
    }
    if (el.empty()) return 0.0;
    auto dp = up(block);
    double t = 0;
    dftc::check(ctx(), dft_kinetic_energy(ctx(), &g, DFT_F64, dp.get(), (int64_t)n, (int)el.size(), el.data(), &t),
                "dft_kinetic_energy");
    return t;
}

double calculateMolecularExternalEnergy(const CartesianGrid& grid, const std::vector<double>& density,
                                        const std::vector<double>& vnuc) {
    if (density.size() != grid.getSize() || vnuc.size() != grid.getSize())
        throw std::invalid_argument("La malla, la densidad y el potencial nuclear deben tener el mismo tamano.");
    auto dr = up(density), dv = up(vnuc);
    double s = 0;
    dftc::check(ctx(), dft_reduce(ctx(), DFT_RED_DOT, DFT_F64, dr.get(), dv.get(), (int64_t)density.size(), &s), "dft_reduce");
    return s * (grid.getDx() * grid.getDy() * grid.getDz());
}

EnergyComponents calculateMolecularTotalEnergy(const XCFunctional& functional, const Molecule& molecule,
                                               const CartesianGrid& grid, const std::vector<double>& na,
                                               const std::vector<double>& nb,
                                               const std::vector<MolecularOrbital>& orbitals,
                                               const std::vector<double>& vh, const std::vector<double>& vnuc) {
    const size_t n = grid.getSize();
    if (na.size() != n || nb.size() != n)
        throw std::invalid_argument("Las densidades moleculares y la malla deben tener el mismo tamano.");
    if (vh.size() != n) throw std::invalid_argument("El potencial de Hartree y la malla deben tener el mismo tamano.");
    if (vnuc.size() != n) throw std::invalid_argument("El potencial nuclear y la malla deben tener el mismo tamano.");
    if (molecule.getNucleusCount() == 0) throw std::invalid_argument("La molecula debe contener al menos un nucleo.");
    std::vector<double> rho(n);
    for (size_t i = 0; i < n; ++i) rho[i] = na[i] + nb[i];
    EnergyComponents e;
    e.kinetic = calculateMolecularKineticEnergy(grid, orbitals);
    e.external = calculateMolecularExternalEnergy(grid, rho, vnuc);
    e.hartree = calculateHartreeEnergy(grid, rho, vh);
    e.exchangeCorrelation = calculateMolecularExchangeCorrelationEnergy(functional, grid, na, nb);
    e.nuclearRepulsion = NuclearRepulsion::calculate(molecule);
    e.total = e.kinetic + e.external + e.hartree + e.exchangeCorrelation + e.nuclearRepulsion;
    return e;
}

// ---- NuclearPotential.cpp / NuclearRepulsion.cpp --------------------------------
std::vector<double> NuclearPotential::calculate(const Molecule& molecule, const CartesianGrid& grid) {
    const dft_grid g = to_grid(grid);
    std::vector<double> nuc;
    for (const auto& a : molecule.getNuclei()) nuc.insert(nuc.end(), {(double)a.atomicNumber, a.position[0], a.position[1], a.position[2]});
    dftc::Buf v(ctx(), grid.getSize() * 8);
    dftc::check(ctx(), dft_nuclear_potential(ctx(), &g, DFT_F64, (int)molecule.getNucleusCount(), nuc.data(), v.get()),
                "dft_nuclear_potential");
    return down(v, grid.getSize());
}

double NuclearPotential::calculateAtPoint(const Molecule& molecule, double x, double y, double z) {
    double p = 0.0;  // NuclearPotential.cpp:36-76 (a single point: no grid work)
    for (const auto& a : molecule.getNuclei()) {
        const double dx = x - a.position[0], dy = y - a.position[1], dz = z - a.position[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 == 0.0) throw std::runtime_error("NuclearPotential is singular at a nuclear position.");
        p -= (double)a.atomicNumber / std::sqrt(d2);
    }
    return p;
}

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

// ---- EigenvalueSolverMolecule.cpp:666-1300 --------------------------------------
std::vector<MolecularOrbital> solveMolecularOrbitals(const CartesianGrid& grid, const std::vector<double>& veff,
                                                     std::size_t count, const std::vector<int>& occupations,
                                                     SpinChannel spin, const std::vector<MolecularOrbital>& init,
                                                     std::size_t maxIterations) {
    if (count == 0) throw std::invalid_argument("El numero de orbitales debe ser mayor que cero.");
    if (occupations.size() != count) throw std::invalid_argument("Las ocupaciones no coinciden con el numero de orbitales.");
    if (!init.empty() && init.size() != count)
        throw std::invalid_argument("Los orbitales iniciales no coinciden con el numero solicitado.");
    for (int o : occupations)
        if (o < 0 || o > 2) throw std::invalid_argument("Ocupacion orbital invalida.");
    if (veff.size() != grid.getSize()) throw std::invalid_argument("El potencial y la malla no coinciden.");
    if (maxIterations == 0) throw std::invalid_argument("maxIterations debe ser mayor que cero.");
    const dft_grid g = to_grid(grid);
    const size_t n = grid.getSize();
    const int nc = (int)count;
    const int k = nc + std::max(2, nc / 4);  // a few buffer states speed up the highest requested one
    dftc::BlockDavidson D(ctx(), g, DFT_F64, std::min<int64_t>(k, (int64_t)n / 3 > nc ? k : nc), 3 * k);
    auto dv = up(veff);
    bool warm = false;
    if (!init.empty()) {
        std::vector<double> block((size_t)nc * n);
        bool finite = true;
        for (int i = 0; i < nc && finite; ++i) {
            if (init[i].psi.size() != n) { finite = false; break; }
            for (double x : init[i].psi) if (!std::isfinite(x)) { finite = false; break; }
            std::copy(init[i].psi.begin(), init[i].psi.end(), block.begin() + (size_t)i * n);
        }
        if (finite) {  // EigenvalueSolverMolecule.cpp:592-624: unusable guesses fall back to seeds
            auto db = up(block);
            D.load_X(db.get(), DFT_F64, nc);
            D.seed_columns(nc);
            warm = true;
        }
    }
    dftc::DiagonalPreconditioner P;
    dftc::DavidsonOptions o;
    o.nconv = nc;
    o.tol = 1.0e-7;
    o.max_iter = (int)std::max<std::size_t>(1, maxIterations / 2);
    dftc::DavidsonStats st = D.solve(dv.get(), warm, o, P);
    if (!st.converged)
        throw std::runtime_error("El orbital molecular no convergio. Orbital = " + std::to_string(nc - 1) +
                                 ", residual = " + std::to_string(st.max_residual));
    dftc::Buf tmp(ctx(), (size_t)nc * n * 8);
    dftc::check(ctx(), dft_copy(ctx(), tmp.get(), D.X(), (size_t)nc * n * 8), "dft_copy");
    std::vector<double> psi = down(tmp, (size_t)nc * n);
    const double s = 1.0 / std::sqrt(grid.getDx() * grid.getDy() * grid.getDz());
    std::vector<MolecularOrbital> out(nc);
    for (int i = 0; i < nc; ++i) {
        out[i].spin = spin;
        out[i].electrons = occupations[i];
        out[i].eigenvalue = D.eigenvalues()[i];
        out[i].psi.assign(psi.begin() + (size_t)i * n, psi.begin() + (size_t)(i + 1) * n);
        for (double& x : out[i].psi) x *= s;
    }
    return out;
}

MolecularOrbital solveMolecularOrbital(const CartesianGrid& grid, const std::vector<double>& veff, std::size_t index,
                                       SpinChannel spin, int electrons, const std::vector<MolecularOrbital>& previous,
                                       const MolecularOrbital* guess, std::size_t maxIterations) {
    // The block solver returns the lowest index+1 eigenpairs; for exact
    // previous eigenvectors that is the reference's deflated solve.
    if (electrons < 0 || electrons > 2) throw std::invalid_argument("Ocupacion orbital invalida.");
    const std::size_t count = std::max<std::size_t>(index, previous.size()) + 1;
    std::vector<int> occ(count, 1);
    occ.back() = electrons;
    std::vector<MolecularOrbital> init;
    if (guess && previous.size() + 1 == count) {
        init = previous;
        init.push_back(*guess);
    }
    auto all = solveMolecularOrbitals(grid, veff, count, occ, spin, init, maxIterations);
    return all.back();
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
