/*
 * dft_molecule.h -- the reference's C++ solver interface, served by the GPU.
 *
 * Same names, signatures, value semantics and exception types as the
 * reference's molecular path (willandru/DFT src/dft), so host code written
 * against those headers -- the SCF driver in SelfConsistentFieldMolecule.cpp
 * and MolecularSCFIteration.cpp -- compiles against this one header and
 * libdft_core.so.  Every grid function runs on the GPU through
 * include/dft_hip.h (one context per host thread, device 0 unless
 * dft::setDevice was called); there is no CPU implementation behind it.
 * Differences from the reference, by design:
 *   - solveMolecularOrbitals solves the block of states at once (block
 *     Davidson with MFMA Gram products) instead of one orbital at a time; it
 *     returns the same eigenpairs to the residual tolerance 1e-7.
 *   - XC functionals are dispatched to the device by their dynamic type
 *     (PZ81, PBE96); any other XCFunctional subclass is rejected with
 *     std::invalid_argument.
 * The types mirror the reference's DFTData.h, DFTConstants.h, XCInput.h and
 * XCFunctional.h member for member.  This header stands in for the
 * reference's headers; it is not meant to be included beside them.  Inside
 * the reference's own tree the hot-path functions come from
 * dft_amd/host/ks_path.cpp compiled against the reference's headers
 * (-DDFT_REFERENCE_HEADERS), beneath its unchanged SCF driver.
 */
#pragma once

#include <array>
#include <chrono>
#include <cstddef>
#include <limits>
#include <string>
#include <vector>

// ---- DFTConstants.h:5-27 --------------------------------------------------
namespace DFTConstants {
constexpr double PI = 3.141592653589793238462643383279502884;
constexpr double EPS = 1.0e-14;
constexpr double RHO_FLOOR = 1.0e-20;
constexpr std::size_t GRID_POINTS = 2000;
constexpr double RMAX = 30.0;
constexpr double MIXING = 0.25;
constexpr int MAX_SCF_ITERATIONS = 250;
constexpr double DENSITY_TOL = 1.0e-9;
constexpr double ENERGY_TOL = 1.0e-11;
constexpr double KS_RESIDUAL_TOL = 1.0e-7;
constexpr double H_EXACT = -0.5;
constexpr double ANGULAR_MOMENTUM_EPS = 1.0e-14;
constexpr double EXCHANGE_COEFFICIENT = 0.7385587663820223;
}  // namespace DFTConstants

// ---- CartesianGrid.h:5-61 ---------------------------------------------------
class CartesianGrid {
   public:
    CartesianGrid(std::size_t nx, std::size_t ny, std::size_t nz, double xmin, double xmax, double ymin, double ymax,
                  double zmin, double zmax);
    std::size_t getNx() const { return nx_; }
    std::size_t getNy() const { return ny_; }
    std::size_t getNz() const { return nz_; }
    std::size_t getSize() const { return nx_ * ny_ * nz_; }
    double getXMin() const { return xmin_; }
    double getXMax() const { return xmax_; }
    double getYMin() const { return ymin_; }
    double getYMax() const { return ymax_; }
    double getZMin() const { return zmin_; }
    double getZMax() const { return zmax_; }
    double getDx() const { return dx_; }
    double getDy() const { return dy_; }
    double getDz() const { return dz_; }
    double getX(std::size_t i) const;
    double getY(std::size_t j) const;
    double getZ(std::size_t k) const;
    std::size_t getIndex(std::size_t i, std::size_t j, std::size_t k) const;

   private:
    std::size_t nx_, ny_, nz_;
    double xmin_, xmax_, ymin_, ymax_, zmin_, zmax_;
    double dx_, dy_, dz_;
};

// ---- Molecule.h:5-36 ---------------------------------------------------------
class Molecule {
   public:
    struct Nucleus {
        int atomicNumber;
        std::array<double, 3> position;
    };
    explicit Molecule(int charge = 0) : charge_(charge) {}
    void addNucleus(int atomicNumber, double x, double y, double z);
    void clear() { nuclei_.clear(); }
    void setCharge(int charge) { charge_ = charge; }
    int getCharge() const { return charge_; }
    int getElectronCount() const;
    std::size_t getNucleusCount() const { return nuclei_.size(); }
    const std::vector<Nucleus>& getNuclei() const { return nuclei_; }

   private:
    std::vector<Nucleus> nuclei_;
    int charge_;
};

// ---- DFTData.h:7-91 ------------------------------------------------------------
enum class SpinChannel { Alpha, Beta };

struct ElectronicState {
    int n = 0;
    int l = 0;
    int alphaElectrons = 0;
    int betaElectrons = 0;
};

struct AtomicConfiguration {
    int Z = 0;
    std::string symbol;
    std::vector<ElectronicState> states;
};

struct AtomicOrbital {
    int n = 0;
    int l = 0;
    SpinChannel spin = SpinChannel::Alpha;
    int electrons = 0;
    double eigenvalue = 0.0;
    std::vector<double> u;
};

struct MolecularOrbital {
    SpinChannel spin = SpinChannel::Alpha;
    int electrons = 0;
    double eigenvalue = 0.0;
    std::vector<double> psi;
};

struct TridiagonalMatrix {
    std::vector<double> lower;
    std::vector<double> diagonal;
    std::vector<double> upper;
};

struct EnergyComponents {
    double kinetic = 0.0;
    double external = 0.0;
    double hartree = 0.0;
    double exchangeCorrelation = 0.0;
    double nuclearRepulsion = 0.0;
    double total = 0.0;
};

struct SCFResult {
    std::vector<AtomicOrbital> orbitals;  // radial solver only; the molecular path leaves it empty
    std::vector<MolecularOrbital> molecularOrbitals;
    std::vector<double> alphaDensity;
    std::vector<double> betaDensity;
    std::vector<double> density;
    std::vector<double> alphaEffectivePotential;
    std::vector<double> betaEffectivePotential;
    std::vector<double> effectivePotential;
    EnergyComponents energy;
    double densityDifference = 0.0;
    double energyDifference = 0.0;
    double maxKSResidual = 0.0;
    int iterations = 0;
    bool converged = false;
};

struct AtomicResult {
    int Z = 0;
    std::string symbol;
    int electrons = 0;
    SCFResult scf;
};

struct MolecularResult {
    int electrons = 0;
    int charge = 0;
    SCFResult scf;
};

// ---- XCInput.h, XCFunctional.h, PZ81.h, PBE96.h ------------------------------
struct XCInput {
    double alphaDensity;
    double betaDensity;
    double alphaGradient;
    double betaGradient;
};

struct XCResult {
    double energyPerElectron;
    double potentialAlpha;
    double potentialBeta;
    double gradientCoefficientAlpha;
    double gradientCoefficientBeta;
};

class XCFunctional {
   public:
    virtual ~XCFunctional() = default;
    virtual XCResult evaluate(const XCInput& input) const = 0;
};

class PZ81 final : public XCFunctional {
   public:
    XCResult evaluate(const XCInput& input) const override;
};

class PBE96 final : public XCFunctional {
   public:
    XCResult evaluate(const XCInput& input) const override;
};

// ---- grid kernels (KohnShamHamiltonianMolecule.h, HartreePotentialMolecule.h,
//      ExchangeCorrelationMolecule.h, ElectronDensityMolecule.h, TotalEnergyMolecule.h)
std::vector<double> applyMolecularKohnShamHamiltonian(const CartesianGrid& grid,
                                                      const std::vector<double>& effectivePotential,
                                                      const std::vector<double>& psi);

std::vector<double> calculateHartreePotential(const CartesianGrid& grid, const std::vector<double>& density);
double calculateHartreeEnergy(const CartesianGrid& grid, const std::vector<double>& density,
                              const std::vector<double>& hartreePotential);

std::vector<double> calculateMolecularSpinExchangeCorrelationPotential(const XCFunctional& functional,
                                                                       const CartesianGrid& grid,
                                                                       const std::vector<double>& alphaDensity,
                                                                       const std::vector<double>& betaDensity,
                                                                       int spin);
double calculateMolecularExchangeCorrelationEnergy(const XCFunctional& functional, const CartesianGrid& grid,
                                                   const std::vector<double>& alphaDensity,
                                                   const std::vector<double>& betaDensity);

std::vector<double> calculateMolecularSpinDensity(const std::vector<MolecularOrbital>& orbitals, SpinChannel spin);
// n = n_alpha + n_beta (ElectronDensityMolecule.cpp:78-114); runtime_error when the two differ in length
std::vector<double> calculateMolecularElectronDensity(const std::vector<MolecularOrbital>& orbitals);
// sum(n) dx dy dz (ElectronDensityMolecule.cpp:116-140)
double integrateMolecularElectronDensity(const std::vector<double>& density, double dx, double dy, double dz);

double calculateMolecularKineticEnergy(const CartesianGrid& grid, const std::vector<MolecularOrbital>& orbitals);
double calculateMolecularExternalEnergy(const CartesianGrid& grid, const std::vector<double>& density,
                                        const std::vector<double>& nuclearPotential);
EnergyComponents calculateMolecularTotalEnergy(const XCFunctional& functional, const Molecule& molecule,
                                               const CartesianGrid& grid, const std::vector<double>& alphaDensity,
                                               const std::vector<double>& betaDensity,
                                               const std::vector<MolecularOrbital>& orbitals,
                                               const std::vector<double>& hartreePotential,
                                               const std::vector<double>& nuclearPotential);

class NuclearPotential {
   public:
    static std::vector<double> calculate(const Molecule& molecule, const CartesianGrid& grid);
    static double calculateAtPoint(const Molecule& molecule, double x, double y, double z);
};

class NuclearRepulsion {
   public:
    static double calculate(const Molecule& molecule);
};

// ---- eigensolver (EigenvalueSolverMolecule.h:9-27) ----------------------------
MolecularOrbital solveMolecularOrbital(const CartesianGrid& grid, const std::vector<double>& effectivePotential,
                                       std::size_t orbitalIndex, SpinChannel spin, int electrons,
                                       const std::vector<MolecularOrbital>& previousOrbitals,
                                       const MolecularOrbital* initialGuess = nullptr,
                                       std::size_t maxIterations = 5000);

std::vector<MolecularOrbital> solveMolecularOrbitals(const CartesianGrid& grid,
                                                     const std::vector<double>& effectivePotential,
                                                     std::size_t numberOfOrbitals, const std::vector<int>& occupations,
                                                     SpinChannel spin,
                                                     const std::vector<MolecularOrbital>& initialOrbitals = {},
                                                     std::size_t maxIterations = 5000);

// ---- SCF helpers and driver (MolecularSCFHelpers.h, MolecularSCFIteration.h,
//      SelfConsistentFieldMolecule.h) -------------------------------------------
struct MolecularSCFTiming {
    double initialization = 0.0, density = 0.0, hartree = 0.0, exchangeCorrelationAlpha = 0.0,
           exchangeCorrelationBeta = 0.0, potentialConstruction = 0.0, orbitalAlpha = 0.0, orbitalBeta = 0.0,
           outputDensity = 0.0, outputHartree = 0.0, totalEnergy = 0.0, densityDifference = 0.0, ksResidual = 0.0,
           mixing = 0.0, totalIterations = 0.0;
};

struct MolecularSCFIterationTiming {
    double density = 0.0, hartree = 0.0, exchangeCorrelationAlpha = 0.0, exchangeCorrelationBeta = 0.0,
           potentialConstruction = 0.0, orbitalAlpha = 0.0, orbitalBeta = 0.0, outputDensity = 0.0,
           outputHartree = 0.0, totalEnergy = 0.0, densityDifference = 0.0, ksResidual = 0.0, mixing = 0.0,
           total = 0.0;
};

struct MolecularSCFConvergenceAnalysis {
    int lastEnergyCriterionIteration = 0;
    int lastDensityCriterionIteration = 0;
    int lastKSCriterionIteration = 0;
    double finalEnergyDifference = std::numeric_limits<double>::infinity();
    double finalDensityDifference = std::numeric_limits<double>::infinity();
    double finalKSResidual = std::numeric_limits<double>::infinity();
    double finalMixing = 0.0;
    double minimumEnergyDifference = std::numeric_limits<double>::infinity();
    double minimumDensityDifference = std::numeric_limits<double>::infinity();
    double minimumKSResidual = std::numeric_limits<double>::infinity();
};

double elapsedSeconds(const std::chrono::steady_clock::time_point& start,
                      const std::chrono::steady_clock::time_point& end);
void mixDensity(std::vector<double>& density, const std::vector<double>& output, double mixing);
double calculateMolecularDensityDifference(const CartesianGrid& grid, const std::vector<double>& oldDensity,
                                           const std::vector<double>& newDensity);
double calculateMolecularSpinDensityDifference(const CartesianGrid& grid, const std::vector<double>& oldAlphaDensity,
                                               const std::vector<double>& oldBetaDensity,
                                               const std::vector<double>& newAlphaDensity,
                                               const std::vector<double>& newBetaDensity);
double calculateMaximumMolecularKSResidual(const CartesianGrid& grid, const std::vector<double>& alphaPotential,
                                           const std::vector<double>& betaPotential,
                                           const std::vector<MolecularOrbital>& orbitals);
std::vector<int> buildSpinOccupations(int electronCount);
std::vector<MolecularOrbital> convertOrbitalsToSpin(const std::vector<MolecularOrbital>& orbitals, SpinChannel spin);
std::vector<MolecularOrbital> solveMolecularOrbitalsSilently(const CartesianGrid& grid,
                                                             const std::vector<double>& potential,
                                                             std::size_t orbitalCount,
                                                             const std::vector<int>& occupations, SpinChannel spin,
                                                             const std::vector<MolecularOrbital>& initialOrbitals);
void printSCFHeader();
void printSCFIteration(int iteration, double energy, double energyDifference, double densityDifference,
                       double residual, double mixing, double iterationTime,
                       const std::vector<MolecularOrbital>& alphaOrbitals,
                       const std::vector<MolecularOrbital>& betaOrbitals, double alphaTime, double betaTime);
void printProfiling(const MolecularSCFTiming& timing, double scfTotalTime);
void printConvergenceAnalysis(const MolecularSCFConvergenceAnalysis& analysis);
void buildInitialMolecularDensities(const CartesianGrid& grid, const Molecule& molecule, int alphaElectrons,
                                    int betaElectrons, bool closedShell, std::vector<double>& alphaDensity,
                                    std::vector<double>& betaDensity,
                                    std::vector<MolecularOrbital>& initialAlphaOrbitals,
                                    std::vector<MolecularOrbital>& initialBetaOrbitals);

struct MolecularSCFIterationResult {
    std::vector<double> density;
    std::vector<double> hartreePotential;
    std::vector<double> alphaXCPotential;
    std::vector<double> betaXCPotential;
    std::vector<double> alphaPotential;
    std::vector<double> betaPotential;
    std::vector<MolecularOrbital> alphaOrbitals;
    std::vector<MolecularOrbital> betaOrbitals;
    std::vector<MolecularOrbital> orbitals;
    std::vector<double> outputAlphaDensity;
    std::vector<double> outputBetaDensity;
    std::vector<double> outputDensity;
    std::vector<double> outputHartreePotential;
    EnergyComponents energy;
    double densityDifference = 0.0;
    double spinDensityDifference = 0.0;
    double effectiveDensityDifference = 0.0;
    double residual = 0.0;
};

MolecularSCFIterationResult executeMolecularSCFIteration(
    const CartesianGrid& grid, const Molecule& molecule, const XCFunctional& functional,
    const std::vector<double>& nuclearPotential, const std::vector<double>& alphaDensity,
    const std::vector<double>& betaDensity, int alphaElectrons, int betaElectrons, bool closedShell,
    const std::vector<MolecularOrbital>& previousAlphaOrbitals, const std::vector<MolecularOrbital>& previousBetaOrbitals);

MolecularResult solveMolecularSelfConsistentField(const CartesianGrid& grid, const Molecule& molecule, int charge,
                                                  int multiplicity, const XCFunctional& functional);

namespace dft {
// Select the GPU used by this thread's implicit context (default 0).
void setDevice(int device);
}  // namespace dft
