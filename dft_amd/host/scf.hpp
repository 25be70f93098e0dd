// scf.hpp -- device-resident Kohn-Sham SCF driver.
//
// Restates SelfConsistentFieldMolecule.cpp:50-561 and
// MolecularSCFIteration.cpp:12-252 step for step (densities -> V_H -> V_xc
// -> V_eff -> orbitals -> output densities -> output V_H -> energies ->
// dRho -> KS residual -> stop test -> lagged adaptive linear mixing) with
// every field on the GPU and the orbital solve done by the block Davidson.
// Mixed mode runs the loop in FP32 (split-BF16 Gram products) until the
// stop rule holds, then continues in FP64 until it holds again.
#pragma once

#include <memory>
#include <vector>

#include "davidson.hpp"
#include "dft_core.h"
#include "gpu.hpp"

namespace dftc {

struct ScfOutputs {
    std::vector<double> eig;    // occupied alpha then beta eigenvalues
    std::vector<double> trace;  // 5 per step
    std::vector<double> rho;    // output alpha then beta density (FP64), optional
    bool want_rho = false;
    // final-step fields for the reference API (FP64, host): occupied orbitals
    // in the reference normalisation (sum psi^2 dV = 1), V_eff per spin, V_H
    bool want_fields = false;
    std::vector<double> psi_a, psi_b, veff_a, veff_b;
    // given input densities (one-step mode) and step-field export
    const double* init_rho_a = nullptr;
    const double* init_rho_b = nullptr;
    const dft_scf_fields* export_fields = nullptr;
};

// Fermi-Dirac occupations of one channel (each in [0, 1], summing to nelec);
// returns the chemical potential.  Bisection on mu over a bracket that holds
// every eigenvalue +- 50 kT, to rounding.
double fermi_occupations(const double* eps, int n, double nelec, double kT, double* occ);
// S = -sum [o ln o + (1 - o) ln(1 - o)] (0 ln 0 = 0)
double fermi_entropy(const double* occ, int n);

// Factory for the learned preconditioner (cnn.cpp); null when unavailable.
// box (optional, {x0, y0, z0, bx, by, bz}): the region a base-1 model covers (null: the whole grid).
std::unique_ptr<Preconditioner> make_cnn_preconditioner(dft_ctx* c, const void* model, const dft_grid& g,
                                                        const int64_t* box = nullptr);
// DFT_PRECOND_* -> preconditioner (throws invalid_argument for CNN without a model)
std::unique_ptr<Preconditioner> make_preconditioner(dft_ctx* c, int kind, const void* model, const dft_grid& g,
                                                    int sweeps = 0, double shift = 0.0, double shift_scale = 0.0,
                                                    const int64_t* box = nullptr);
// The learned correction's box for a molecule: the nuclei's bounding box widened by `margin` grid points on
// every side, clipped to the grid ({x0, y0, z0, bx, by, bz}).
void nuclei_box(const dft_grid& g, int natoms, const double* nuclei, int margin, int64_t box[6]);

class ScfDriver {
   public:
    ScfDriver(dft_ctx* c, const dft_grid& g, int natoms, const double* nuclei, const dft_scf_config& cfg);
    void run(dft_scf_report& rep, ScfOutputs& out);
    // Use p (not owned) for every Davidson correction instead of the one cfg selects.
    void set_preconditioner(Preconditioner* p) { ext_precond_ = p; }

   private:
    struct Fields;
    dft_ctx* c_;
    dft_grid g_;
    std::vector<double> nuc_;
    dft_scf_config cfg_;
    Preconditioner* ext_precond_ = nullptr;
};

}  // namespace dftc
