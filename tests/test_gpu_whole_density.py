"""The total electron density and its integral through the reference-shaped
C++ API on the GPU (calculateMolecularElectronDensity,
integrateMolecularElectronDensity; ElectronDensityMolecule.cpp:78-140).  A
program written against include/dft_molecule.h (tests/cpp/density_api_check.cpp,
built by the build step) reads the golden orbitals, and its outputs are
compared with the reference's own (tests/golden): relative 1e-12, the FP64
bar of the other grid kernels.  It also reports the reference's argument
errors, each with the reference's exception type."""
import os
import subprocess

import numpy as np
import pytest

from conftest import grid_from

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_total_density_and_integral_match_reference(tmp_path, goldens):
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "density_api_check")  # built by __graft_entry__.build()
    assert os.path.exists(exe), "tests/cpp/bin/density_api_check is not built"
    og = grid_from(goldens["edens_grid"])
    psi, el, spins = goldens["edens_psi"], goldens["edens_el"], goldens["edens_spins"]
    np.array([*og.n, *og.bounds], dtype=np.float64).tofile(tmp_path / "grid.bin")
    np.concatenate([[psi.shape[0]], el.astype(float), spins.astype(float), psi.ravel()]).astype(np.float64).tofile(
        tmp_path / "orbitals.bin")
    print(f"[density api] grid {og.n}, {psi.shape[0]} orbitals", flush=True)
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    d = np.fromfile(tmp_path / "density.out", dtype=np.float64)
    ref = goldens["edens_density"]
    assert d.shape == ref.shape
    assert float(np.max(np.abs(d - ref)) / np.max(np.abs(ref))) <= 1e-12
    total = float(np.fromfile(tmp_path / "integral.out", dtype=np.float64)[0])
    want = float(goldens["edens_integral"])
    assert abs(total - want) <= 1e-12 * abs(want)
    errs = np.fromfile(tmp_path / "errors.out", dtype=np.float64)
    assert list(errs) == [1.0, 1.0, 1.0, 1.0], errs
