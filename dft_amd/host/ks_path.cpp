// ks_path.cpp -- the reference's 3D hot-path functions served by the GPU:
// H psi, the Cartesian Hartree potential and energy, the molecular XC
// potential and energy, spin and total densities, the energy terms, the
// nuclear potential and the molecular eigensolver.  Each keeps the
// reference's signature, argument checks, messages and exception types; the
// grid arithmetic goes through libdft_hip.so's C-ABI on a per-thread context.
//
// The file compiles in two settings:
//   - inside this library, against include/dft_molecule.h;
//   - inside the reference's own tree with -DDFT_REFERENCE_HEADERS and its
//     src/dft on the include path, in place of KohnShamHamiltonianMolecule.cpp,
//     HartreePotentialMolecule.cpp, ExchangeCorrelationMolecule.cpp,
//     ElectronDensityMolecule.cpp, TotalEnergyMolecule.cpp,
//     EigenvalueSolverMolecule.cpp and NuclearPotential.cpp; the types then come
//     from the reference's headers and its SCF driver runs unchanged above it.
#ifdef DFT_REFERENCE_HEADERS
#include "EigenvalueSolverMolecule.h"
#include "ElectronDensityMolecule.h"
#include "ExchangeCorrelationMolecule.h"
#include "HartreePotentialMolecule.h"
#include "KohnShamHamiltonianMolecule.h"
#include "NuclearPotential.h"
#include "NuclearRepulsion.h"
#include "PBE96.h"
#include "PZ81.h"
#include "TotalEnergyMolecule.h"
#else
#include "dft_molecule.h"
#endif

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>

#include "api_ctx.hpp"
#include "davidson.hpp"

namespace dft {
void setDevice(int device);
}

namespace dftapi {
namespace {
thread_local int t_device = 0;
thread_local std::unique_ptr<dftc::Ctx> t_ctx;
}  // namespace

dft_ctx* ctx() {
    if (!t_ctx) t_ctx.reset(new dftc::Ctx(t_device));
    return t_ctx->get();
}

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
}  // namespace dftapi

namespace dft {
void setDevice(int device) {
    if (dftapi::t_device != device) dftapi::t_ctx.reset();
    dftapi::t_device = device;
}
}  // namespace dft

using dftapi::ctx;
using dftapi::down;
using dftapi::up;

namespace {

template <class G>
dft_grid to_grid(const G& g) { return dftapi::to_grid(g); }

// XCFunctional -> device functional id: the two functionals the kernels
// implement, recognised by their dynamic type (PZ81.h, PBE96.h).
int device_functional(const XCFunctional& f) {
    if (dynamic_cast<const PZ81*>(&f)) return DFT_XC_PZ81;
    if (dynamic_cast<const PBE96*>(&f)) return DFT_XC_PBE96;
    throw std::invalid_argument("XC functional has no device implementation (PZ81 and PBE96 only)");
}

}  // namespace

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
