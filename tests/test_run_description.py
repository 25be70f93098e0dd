"""Run descriptions (dft_amd/run.py): a result document written by the C++
writer (dft_scf_result_json) reads back as the same SCF configuration, knob
for knob and bit for bit, so a run re-runs from its own result.  Host code
only."""
import json

import numpy as np
import pytest

import dft_amd
from dft_amd import run as R

SKIP = {"cnn", "verbose"}  # a pointer and a logging switch: not part of a run's definition


def _config():
    """A dft_scf_config with every knob away from its default."""
    cfg = dft_amd.ScfConfig()
    vals = dict(functional=dft_amd.PBE96, charge=-1, multiplicity=3, precision=dft_amd.SCF_MIXED, nstates=17,
                max_subspace=61, max_iterations=123, poisson_fixed=77, poisson_batch=32,
                poisson_method=dft_amd.POISSON_CHEBYSHEV, poisson_tol=3.3e-9, stop_rule=dft_amd.STOP_DE,
                energy_tol=2.5e-6, init_guess=dft_amd.INIT_HYDROGENIC, davidson_tol=1.0 / 3.0 * 1e-7,
                davidson_max_iter=41, refine_iterations=9, precond=dft_amd.PRECOND_KINETIC_EXACT, precond_sweeps=5,
                precond_shift=0.7, mixer=dft_amd.MIX_PULAY, mix_depth=8, mix_beta=0.35, smearing=0.004,
                ortho_gram=dft_amd.GRAM_SPLIT_BF16, mix_kerker=0.8, precond_shift_scale=0.45)
    for k, v in vals.items():
        setattr(cfg, k, v)
    return cfg, vals


def _document(cfg, grid):
    nuc = np.array([[8.0, 0.0, 0.1, -0.2], [1.0, 1.43, 0.0, 1.108], [1.0, -1.43, 0.0, 1.108]])
    res = dict(config=cfg, report=dft_amd.ScfReport(), eigenvalues=np.array([-1.0, -0.5, -0.9]), n_alpha=2, n_beta=1)
    return nuc, dft_amd.scf_result_json(grid, nuc, res)


def test_every_knob_is_in_the_document():
    covered = {n for n, _ in R._NAMED.values()} | set(R._INTS.values()) | set(R._FLOATS.values()) | {"charge"}
    fields = {k for k, _ in dft_amd.ScfConfig._fields_} - SKIP
    assert covered == fields


def test_result_document_reads_back_as_its_configuration():
    grid = dft_amd.Grid.make(33, 29, 31, -6.4, 6.4, -5.6, 5.6, -6.0, 6.0)
    cfg, vals = _config()
    nuc, text = _document(cfg, grid)
    g2, nuc2, kw = R.parse_run(text)
    assert bytes(g2) == bytes(grid)
    assert np.array_equal(nuc2, nuc)
    for k, _ in dft_amd.ScfConfig._fields_:
        if k in SKIP:
            continue
        assert kw[k] == getattr(cfg, k), k
    assert kw["davidson_tol"] == vals["davidson_tol"]  # 17 digits: the exact double
    # and again through a file, the way `python -m dft_amd.run` reads it
    doc = json.loads(text)
    assert R.parse_run(doc)[2] == kw


def test_minimal_description_takes_the_defaults(tmp_path):
    doc = {"grid": {"n": [7, 7, 7], "min": [-8.1, -8.1, -8.1], "max": [7.9, 7.9, 7.9]},
           "nuclei": [{"Z": 1, "position": [-0.7, 0, 0]}, {"Z": 1, "position": [0.7, 0, 0]}],
           "config": {"functional": "PZ81"}}
    p = tmp_path / "h2.json"
    p.write_text(json.dumps(doc))
    g, nuc, kw = R.parse_run(p)
    assert (g.nx, g.ny, g.nz) == (7, 7, 7) and g.dx == 16.0 / 6
    assert nuc.shape == (2, 4) and kw == {"functional": dft_amd.PZ81}


def test_description_errors():
    base = {"grid": {"n": [5, 5, 5], "min": [-1, -1, -1], "max": [1, 1, 1]}, "nuclei": []}
    with pytest.raises(ValueError, match="unknown config key"):
        R.parse_run({**base, "config": {"energyTol": 1e-5}})
    with pytest.raises(ValueError, match="unknown run keys"):
        R.parse_run({**base, "geometry": []})
    with pytest.raises(ValueError, match="not one of"):
        R.parse_run({**base, "config": {"functional": "B3LYP"}})
    with pytest.raises(ValueError, match="disagrees"):
        R.parse_run({"grid": {**base["grid"], "spacing": [0.5, 0.5, 0.4]}, "nuclei": []})


def test_check_mode_compares_a_rerun(tmp_path, monkeypatch):
    """--check re-runs a result document and passes only when every energy term and the step count repeat."""
    grid = dft_amd.Grid.make(7, 7, 7, -8.1, 7.9, -8.1, 7.9, -8.1, 7.9)
    cfg, _ = _config()
    rep = dft_amd.ScfReport()
    energies = [1.5, -4.0, 2.0, -0.75, 0.7142857142857143, -0.5357142857142857]
    for i, e in enumerate(energies):
        rep.energy[i] = e
    rep.iterations = 12
    nuc = np.array([[1.0, -0.7, 0.0, 0.0], [1.0, 0.7, 0.0, 0.0]])
    res = dict(config=cfg, report=rep, eigenvalues=np.array([-0.3, -0.3]), n_alpha=1, n_beta=1)
    path = tmp_path / "result.json"
    path.write_text(dft_amd.scf_result_json(grid, nuc, res))

    def fake_run(doc, want_density=False, verbose=False):
        g, n, kw = R.parse_run(doc)
        assert kw["smearing"] == cfg.smearing  # the knobs of the document drive the re-run
        shift = fake_run.shift
        out = dict(energy=[e + shift for e in energies], total=energies[5] + shift, iterations=12, converged=True,
                   config=cfg, report=rep, eigenvalues=np.array([-0.3, -0.3]), n_alpha=1, n_beta=1)
        return g, n, out

    monkeypatch.setattr(R, "run", fake_run)
    fake_run.shift = 0.0
    assert R.main([str(path), "--out", str(tmp_path / "rerun.json"), "--check", "1e-12"]) == 0
    fake_run.shift = 1e-9
    assert R.main([str(path), "--out", str(tmp_path / "rerun.json"), "--check", "1e-10"]) == 3
    assert R.main([str(path), "--out", str(tmp_path / "rerun.json"), "--check", "1e-8"]) == 0
