"""The drop-in boundary, built on the CPU: the reference's SCF driver files and its main.cpp compile
unchanged against the reference's own headers with dft_amd/host/ks_path.cpp in place of the seven
hot-path sources (KohnShamHamiltonianMolecule.cpp, HartreePotentialMolecule.cpp,
ExchangeCorrelationMolecule.cpp, ElectronDensityMolecule.cpp, TotalEnergyMolecule.cpp,
EigenvalueSolverMolecule.cpp, NuclearPotential.cpp), and link to libdft_hip.so.  Skipped where the
reference checkout is absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src/dft"

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout absent")

HOT_PATH_SYMBOLS = ["applyMolecularKohnShamHamiltonian", "calculateMolecularSpinExchangeCorrelationPotential",
                    "calculateMolecularExchangeCorrelationEnergy", "calculateMolecularSpinDensity",
                    "calculateMolecularKineticEnergy", "calculateMolecularTotalEnergy", "solveMolecularOrbitals",
                    "solveMolecularOrbital", "NuclearPotential9calculate"]


def _built(target, exe):
    if not os.path.exists(os.path.join(ROOT, "dft_amd", "lib", "libdft_hip.so")):
        pytest.skip("libdft_hip.so not built")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), target], check=True)
    path = os.path.join(ROOT, "oracle", "_ref", exe)
    assert os.path.exists(path)
    return path


@pytest.mark.parametrize("target,exe", [("dropin", "dropin_scf"), ("dropin-main", "main_dropin")])
def test_reference_driver_links_over_the_gpu_path(target, exe):
    path = _built(target, exe)
    ldd = subprocess.run(["ldd", path], capture_output=True, text=True).stdout
    assert "libdft_hip.so" in ldd
    syms = subprocess.run(["nm", "-C", "--defined-only", path], capture_output=True, text=True).stdout
    # the hot-path functions are defined once, by ks_path.cpp
    for s in HOT_PATH_SYMBOLS:
        assert s.replace("9", "::") in syms or s in syms, s
    # and the reference driver is the reference's own
    assert "executeMolecularSCFIteration" in syms and "solveMolecularSelfConsistentField" in syms


def test_hot_path_sources_left_out():
    """The recipe compiles every reference source except the seven hot-path files."""
    mk = open(os.path.join(ROOT, "oracle", "Makefile")).read()
    for f in ("KohnShamHamiltonianMolecule", "HartreePotentialMolecule", "ExchangeCorrelationMolecule",
              "ElectronDensityMolecule", "TotalEnergyMolecule", "EigenvalueSolverMolecule", "NuclearPotential"):
        assert f in mk.split("HOT_PATH :=")[1].split("DRIVER_SRCS")[0]
