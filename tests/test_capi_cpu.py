"""CPU checks of the native boundary: both libraries load, export every
function include/*.h declares, and the host-only helpers behave (no GPU
compute is called here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|void\*|const char\*|void)\s+\**(dft_\w+)\s*\(", text)))


def exported(lib):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


@pytest.fixture(scope="module")
def libs():
    import dft_amd
    return dft_amd.hip_lib(), dft_amd.core_lib()


def test_hip_library_exports_header(libs):
    import dft_amd
    names = declared("dft_hip.h")
    assert len(names) > 40
    missing = sorted(set(names) - exported(dft_amd.HIP_SO))
    assert not missing, missing


def test_core_library_exports_header(libs):
    import dft_amd
    names = declared("dft_core.h")
    assert {"dft_scf_run", "dft_davidson_run", "dft_core_last_error"} <= set(names)
    missing = sorted(set(names) - exported(dft_amd.CORE_SO))
    assert not missing, missing


def test_reference_cpp_interface_exported(libs):
    """The drop-in C++ names of the reference (mangled) are defined in libdft_core.so."""
    import subprocess

    import dft_amd
    out = subprocess.run(["nm", "-DC", "--defined-only", dft_amd.CORE_SO], capture_output=True, text=True).stdout
    for sym in ["applyMolecularKohnShamHamiltonian(CartesianGrid const&", "calculateHartreePotential(CartesianGrid const&",
                "calculateMolecularSpinExchangeCorrelationPotential(XCFunctional const&",
                "calculateMolecularSpinDensity(std::vector<MolecularOrbital", "solveMolecularOrbitals(CartesianGrid const&",
                "solveMolecularOrbital(CartesianGrid const&", "calculateMolecularTotalEnergy(XCFunctional const&",
                "executeMolecularSCFIteration(CartesianGrid const&", "solveMolecularSelfConsistentField(CartesianGrid const&",
                "NuclearPotential::calculate(Molecule const&", "NuclearRepulsion::calculate(Molecule const&"]:
        assert sym in out, sym


def test_grid_init_matches_reference_spacing(libs):
    import dft_amd
    g = dft_amd.Grid.make(9, 8, 7, -3.1, 2.9, -2.0, 2.5, -1.7, 2.2)
    assert g.dx == (2.9 - -3.1) / 8 and g.dy == (2.5 - -2.0) / 7 and g.dz == (2.2 - -1.7) / 6
    bad = dft_amd.Grid()
    assert libs[0].dft_grid_init(C.byref(bad), 1, 5, 5, *[C.c_double(v) for v in (0, 1, 0, 1, 0, 1)]) == -1
    assert libs[0].dft_grid_init(C.byref(bad), 5, 5, 5, *[C.c_double(v) for v in (1, 0, 0, 1, 0, 1)]) == -1


def test_null_context_is_rejected(libs):
    L = libs[0]
    assert L.dft_sync(None) == -1
    assert L.dft_free(None, None) == -1
    L.dft_last_error.restype = C.c_char_p
    assert L.dft_last_error(None) == b"null context"


def test_no_gpu_context_creation_fails_cleanly(libs):
    import torch  # noqa: F401  (only to know whether a GPU is visible)
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    ctx = C.c_void_p()
    assert libs[0].dft_ctx_create(0, C.byref(ctx)) == -3
