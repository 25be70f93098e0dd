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
    return sorted(set(re.findall(r"\b(?:int|void\*|const char\*|void|uint32_t)\s+\**(dft_\w+)\s*\(", text)))


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


def test_io_library_exports_header(libs):
    import dft_amd
    names = declared("dft_io.h")
    assert {"dft_crc32", "dft_field_write", "dft_field_read", "dft_patch_shard_write", "dft_patch_shard_read",
            "dft_scf_result_json"} <= set(names)
    missing = sorted(set(names) - exported(dft_amd.CORE_SO))
    assert not missing, missing


@pytest.mark.parametrize("header", ["dft_hip.h", "dft_core.h", "dft_io.h"])
def test_c_abi_headers_are_plain_c(header):
    """The C-ABI headers compile as strict C99 (no C++ in them), so C and FFI callers can include them."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), "-x", "c", "-"], input=f'#include "{header}"\n',
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr


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


def test_fermi_occupations(libs):
    """Host occupation solver of the smeared SCF (no GPU): the electron count,
    the chemical potential against an independent root find, and cases the
    mathematics fixes: a degenerate level shares its electrons equally, a
    particle-hole symmetric spectrum at half filling puts mu at its centre,
    a state at mu holds 1/2, a wide gap gives whole orbitals, and the entropy
    of k equal occupations is k times the binary entropy."""
    import numpy as np
    from scipy.optimize import brentq
    import dft_amd
    rng = np.random.default_rng(3)
    eps = np.sort(rng.uniform(-2.0, 0.5, 40))
    for nel, kT in ((17.0, 0.01), (3.5, 0.2), (39.2, 0.05)):
        occ, mu, S = dft_amd.fermi_occupations(eps, nel, kT)
        assert abs(occ.sum() - nel) <= 1e-11
        assert np.all(np.diff(occ) <= 1e-15) and np.all((occ >= 0) & (occ <= 1))
        mu_ref = brentq(lambda m: np.sum(0.5 * (1.0 - np.tanh((eps - m) / (2 * kT)))) - nel, -10, 10, xtol=1e-15)
        assert abs(mu - mu_ref) <= 1e-9
    # degenerate triple holding 1 electron above a gap, below a gap
    e3 = np.array([-5.0, -5.0, -1.0, -1.0, -1.0, 2.0, 3.0])
    occ, mu, S = dft_amd.fermi_occupations(e3, 3.0, 0.01)
    assert np.allclose(occ[:2], 1.0, atol=1e-12) and np.allclose(occ[2:5], 1.0 / 3.0, atol=1e-12)
    assert occ[5] < 1e-100
    h = -(1 / 3 * np.log(1 / 3) + 2 / 3 * np.log(2 / 3))
    assert abs(S - 3 * h) <= 1e-12
    # symmetric spectrum at half filling: mu at 0, the state at mu holds 1/2
    es = np.array([-1.0, -0.3, 0.0, 0.3, 1.0])
    occ, mu, S = dft_amd.fermi_occupations(es, 2.5, 0.1)
    assert abs(mu) <= 1e-12 and abs(occ[2] - 0.5) <= 1e-12 and abs(occ[0] + occ[4] - 1.0) <= 1e-12
    # a wide gap: whole orbitals, no entropy
    occ, mu, S = dft_amd.fermi_occupations(np.array([-3.0, -2.0, 1.0, 2.0]), 2.0, 0.01)
    assert np.array_equal(occ[:2], [1.0, 1.0]) and occ[2] < 1e-60 and S < 1e-60  # mid-gap mu: e^-150
    # full and empty channels; bad arguments raise
    assert np.all(dft_amd.fermi_occupations(es, 5.0, 0.1)[0] == 1.0)
    assert np.all(dft_amd.fermi_occupations(es, 0.0, 0.1)[0] == 0.0)
    with pytest.raises(dft_amd.DftError):
        dft_amd.fermi_occupations(es, 6.0, 0.1)
    with pytest.raises(dft_amd.DftError):
        dft_amd.fermi_occupations(es, 1.0, 0.0)
