"""Run an SCF from a run description and write its result document.

The run description is the result document's own input part (include/dft_io.h):
a JSON object with "grid", "nuclei", "charge" and "config", whose keys are
the ones dft_scf_result_json writes.  A result document is therefore also a
run description: re-running it repeats the run with every knob as it was.
Keys of the result part ("scf", "timing", "counters", ...) are ignored; an
unknown key inside "config" is an error, so a misspelt knob never falls back
to its default silently.  Keys left out take scf_run's defaults.

The reference builds its molecules in main.cpp and prints its results; it
reads no run description.

  python -m dft_amd.run run.json --out result.json [--fields]
  python -m dft_amd.run result.json --out rerun.json --check 1e-10   # repeat a result and compare
"""
import argparse
import json
import os
import sys

import numpy as np

from . import (INIT_BARE, INIT_HYDROGENIC, MIX_LINEAR, MIX_PULAY, PBE96, POISSON_CHEBYSHEV, POISSON_DIRECT,
               POISSON_JACOBI, PRECOND_CNN, PRECOND_DIAG, PRECOND_KINETIC, PRECOND_KINETIC_CHEB, PRECOND_KINETIC_EXACT,
               PZ81, SCF_FP64, SCF_MIXED, STOP_DE, STOP_REFERENCE, CnnModel, Grid, save_scf_result, scf_run)

_NAMED = {
    "functional": ("functional", {"PZ81": PZ81, "PBE96": PBE96}),
    "precision": ("precision", {"fp64": SCF_FP64, "mixed": SCF_MIXED}),
    "poisson": ("poisson_method", {"jacobi": POISSON_JACOBI, "chebyshev": POISSON_CHEBYSHEV, "direct": POISSON_DIRECT}),
    "stopRule": ("stop_rule", {"reference": STOP_REFERENCE, "energy": STOP_DE}),
    "initialGuess": ("init_guess", {"bare": INIT_BARE, "hydrogenic": INIT_HYDROGENIC}),
    "preconditioner": ("precond", {"diagonal": PRECOND_DIAG, "cnn": PRECOND_CNN, "kinetic": PRECOND_KINETIC,
                                   "kinetic_chebyshev": PRECOND_KINETIC_CHEB, "kinetic_exact": PRECOND_KINETIC_EXACT}),
    "mixer": ("mixer", {"linear": MIX_LINEAR, "pulay": MIX_PULAY}),
    "orthoGram": ("ortho_gram", {"default": 0, "f32": 1, "split_bf16": 2}),
    "xcFlux": ("xc_flux", {"reference": 0, "consistent": 1}),
}
_INTS = {"multiplicity": "multiplicity", "nstates": "nstates", "poissonFixedSweeps": "poisson_fixed",
         "maxIterations": "max_iterations", "maxSubspace": "max_subspace", "poissonBatch": "poisson_batch",
         "davidsonMaxIterations": "davidson_max_iter", "refineIterations": "refine_iterations",
         "precondSweeps": "precond_sweeps", "mixDepth": "mix_depth"}
_FLOATS = {"energyTolerance": "energy_tol", "davidsonTolerance": "davidson_tol", "smearing": "smearing",
           "poissonTolerance": "poisson_tol", "precondShift": "precond_shift",
           "precondShiftScale": "precond_shift_scale", "mixBeta": "mix_beta", "mixKerker": "mix_kerker"}
# input-only: the trained network of the learned preconditioner (a CnnModel .npz)
_PATHS = {"cnnModel"}
_TOP = {"grid", "nuclei", "charge", "config"}
_RESULT_PART = {"format", "version", "electrons", "scf", "smearing", "timing", "counters"}


def parse_run(doc):
    """Run description (dict, JSON text or path) -> (grid, nuclei (natoms x 4), scf_run keyword arguments)."""
    if isinstance(doc, (str, os.PathLike)) and not str(doc).lstrip().startswith("{"):
        with open(doc) as f:
            doc = json.load(f)
    elif isinstance(doc, str):
        doc = json.loads(doc)
    unknown = set(doc) - _TOP - _RESULT_PART
    if unknown:
        raise ValueError(f"unknown run keys {sorted(unknown)}")
    gd = doc["grid"]
    n, lo, hi = gd["n"], gd["min"], gd["max"]
    grid = Grid.make(int(n[0]), int(n[1]), int(n[2]), float(lo[0]), float(hi[0]), float(lo[1]), float(hi[1]),
                     float(lo[2]), float(hi[2]))
    if "spacing" in gd:
        got = (grid.dx, grid.dy, grid.dz)
        if any(abs(a - b) > 1e-12 * abs(a) for a, b in zip(got, gd["spacing"])):
            raise ValueError(f"grid spacing {gd['spacing']} disagrees with n, min and max ({got})")
    nuclei = np.array([[float(a["Z"]), *map(float, a["position"])] for a in doc["nuclei"]], dtype=float).reshape(-1, 4)
    kw = {}
    if "charge" in doc:
        kw["charge"] = int(doc["charge"])
    for key, val in doc.get("config", {}).items():
        if key in _NAMED:
            name, table = _NAMED[key]
            if isinstance(val, str):
                if val not in table:
                    raise ValueError(f"config {key}: {val!r} is not one of {sorted(table)}")
                kw[name] = table[val]
            else:
                kw[name] = int(val)  # the document writes a value without a name as its number
        elif key in _INTS:
            kw[_INTS[key]] = int(val)
        elif key in _FLOATS:
            kw[_FLOATS[key]] = float("nan") if val is None else float(val)
        elif key in _PATHS:
            kw["cnn"] = CnnModel.load(val)
        else:
            raise ValueError(f"unknown config key {key!r}")
    return grid, nuclei, kw


def run(doc, want_density=False, verbose=False):
    """Parse a run description and run it; returns (grid, nuclei, scf_run result)."""
    grid, nuclei, kw = parse_run(doc)
    return grid, nuclei, scf_run(grid, nuclei, want_density=want_density, verbose=verbose, **kw)


ENERGY_KEYS = ("kinetic", "external", "hartree", "exchangeCorrelation", "nuclearRepulsion", "total")


def compare_energies(recorded, result, tol):
    """Largest |difference| (Ha) between a result document's energies and a new run's, and whether it is within
    tol; the step counts must agree too."""
    old = recorded["scf"]["energy"]
    # a term recorded as null (non-finite) is never reproduced
    diffs = [abs((float("nan") if old[k] is None else float(old[k])) - float(result["energy"][i]))
             for i, k in enumerate(ENERGY_KEYS)]
    diff = max(diffs) if all(d == d for d in diffs) else float("inf")
    same_steps = int(recorded["scf"]["iterations"]) == int(result["iterations"])
    return diff, diff <= tol and same_steps


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m dft_amd.run", description=__doc__.split("\n\n")[0])
    ap.add_argument("description", help="run description or result document (JSON)")
    ap.add_argument("--out", required=True, help="result document to write")
    ap.add_argument("--fields", action="store_true", help="also write the output spin densities as field files")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--check", type=float, metavar="TOL", default=None,
                    help="the description is a result document: fail unless every energy term is reproduced "
                         "within TOL Ha in the same number of SCF steps")
    a = ap.parse_args(argv)
    recorded = None
    if a.check is not None:
        with open(a.description) as f:
            recorded = json.load(f)
        if "scf" not in recorded:
            ap.error("--check needs a result document (with its \"scf\" section)")
    grid, nuclei, r = run(recorded if recorded is not None else a.description, want_density=a.fields,
                          verbose=a.verbose)
    save_scf_result(a.out, grid, nuclei, r, fields=a.fields)
    print(f"E = {r['total']:.10f} Ha, {r['iterations']} SCF steps, converged {r['converged']}", file=sys.stderr)
    if recorded is not None:
        diff, ok = compare_energies(recorded, r, a.check)
        print(f"check: largest energy difference {diff:.3e} Ha, steps {r['iterations']} vs "
              f"{recorded['scf']['iterations']}: {'reproduced' if ok else 'NOT reproduced'}", file=sys.stderr)
        return 0 if ok else 3
    return 0 if r["converged"] else 1


if __name__ == "__main__":
    sys.exit(main())
