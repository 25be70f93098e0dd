"""The reference's own SCF driver (SelfConsistentFieldMolecule.cpp, MolecularSCFIteration.cpp,
MolecularSCFHelpers.cpp, compiled unchanged) running over this build's GPU hot path
(oracle/_ref/dropin_scf, built by oracle/Makefile where the reference checkout exists).  The total
energies must match main.cpp's known answers (H2 on 7^3, H2O on 9^3, PZ81) and the reference-written
He PBE96 SCF golden within 1e-7 Ha.  Skipped when the program was not built."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_scf")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/dropin_scf not built")]


def run(which):
    print(f"[dropin] {which}", flush=True)
    out = subprocess.run([EXE, which], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    name, e, it, conv = out.stdout.split()
    return float(e), int(it), int(conv)


@pytest.mark.parametrize("which,known", [("h2", -2.8564865005), ("h2o", -101.39828181)])
def test_reference_driver_on_gpu_main_cpp_known_answers(which, known):
    e, it, conv = run(which)
    print(which, e, it, conv)
    assert conv == 1
    assert abs(e - known) < 1e-7


def test_reference_driver_on_gpu_pbe96_golden(goldens):
    e, it, conv = run("he_pbe")
    ref = goldens["scf_he_pbe_result"]
    print("he_pbe", e, it, ref)
    assert conv == 1
    assert abs(e - float(ref[6])) < 1e-7  # fields: iterations, T, Eext, EH, Exc, Enn, total, ...
