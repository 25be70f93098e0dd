"""The learned preconditioner's workload end to end: harvest training pairs
from seeded random molecules into patch shards, train the network on them,
and evaluate it on held-out molecules against the correction it builds on.

Two bases (--base): "kinetic" (default) -- the network learns what to add to
the exact kinetic correction (harvest target_scale 2, model base 1);
"diagonal" -- the network's output is the whole correction (target_scale 1,
model base 0), judged against the diagonal correction.

Defaults are the molecule set of the learned-preconditioner workload: 64
training seeds (1-64) and 8 held-out seeds (1001-1008), 4-12 atoms with Z in
{H, C, N, O} inside a 5-bohr-radius ball (geometry.training_molecule),
spin-unpolarised LSDA/PZ81 on a 128^3 grid with h = 0.2 bohr, 32^3 patches
(about 200k in all), Adam at lr 1e-3 for 20k steps.

  python -m dft_amd.cnn_pipeline harvest  --dir runs/cnn            # one shard per seed, resumable
  python -m dft_amd.cnn_pipeline train    --dir runs/cnn            # runs/cnn/model.npz, train_log.json
  python -m dft_amd.cnn_pipeline evaluate --dir runs/cnn            # runs/cnn/eval.json

The reference has no learned part; its Davidson preconditions with the
diagonal correction only (EigenvalueSolverMolecule.cpp:483-555), which stays
the per-column fallback and the judge of every learned correction.  Every
step runs on the GPU through libdft_core.so; a step that needs the GPU fails
with the library's error when there is none."""
import argparse
import glob
import json
import os
import sys
import time

import numpy as np

from . import (CNN_BASE_KINETIC, CNN_BASE_NONE, MIX_PULAY, CnnModel, CnnTrainer, Grid, INIT_HYDROGENIC,
               POISSON_DIRECT, PRECOND_CNN, PRECOND_DIAG, PRECOND_KINETIC_EXACT, PZ81, SCF_FP64, STOP_DE,
               patch_shard_info, scf_result_json, scf_run)
from .geometry import training_molecule

TRAIN_SEEDS = "1-64"
EVAL_SEEDS = "1001-1008"
# base -> (harvest target_scale, model base, the correction the learned one is judged against)
BASES = {"kinetic": (2, CNN_BASE_KINETIC, PRECOND_KINETIC_EXACT), "diagonal": (1, CNN_BASE_NONE, PRECOND_DIAG)}


def parse_seeds(text):
    """'1-64,70,80-82' -> [1, ..., 64, 70, 80, 81, 82] (order kept, duplicates dropped)."""
    out = []
    for part in str(text).split(","):
        part = part.strip()
        if not part:
            continue
        lo, _, hi = part.partition("-")
        lo, hi = int(lo), int(hi) if hi else int(lo)
        if lo < 0 or hi < lo:
            raise ValueError(f"bad seed range {part!r}")
        out.extend(range(lo, hi + 1))
    seen = set()
    return [s for s in out if not (s in seen or seen.add(s))]


def shard_path(directory, seed):
    return os.path.join(directory, f"seed{seed:05d}.dftp")


def shard_plan(directory):
    """Shards of a harvest directory in seed order, checked for one patch edge.
    The last shard's molecule is held out of training as the validation set.
    Returns dict(paths, counts, patch, total, train_count, val_count)."""
    paths = sorted(glob.glob(os.path.join(directory, "seed*.dftp")))
    if not paths:
        raise FileNotFoundError(f"no patch shards (seed*.dftp) in {directory}")
    heads = [patch_shard_info(p) for p in paths]
    edges = {h["patch"] for h in heads}
    if len(edges) != 1:
        raise ValueError(f"shards hold different patch edges {sorted(edges)}")
    spacings = {h["h"] for h in heads}
    if len(spacings) != 1:
        raise ValueError(f"shards were harvested on different grid spacings {sorted(spacings)}")
    targets = {h["target_scale"] for h in heads}
    if len(targets) != 1:
        raise ValueError(f"shards hold different targets {sorted(targets)}")
    counts = [h["count"] for h in heads]
    total = int(sum(counts))
    val = counts[-1] if len(paths) > 1 else 0
    if total - val < 1:
        raise ValueError("no training patches")
    return dict(paths=paths, counts=counts, patch=edges.pop(), total=total, train_count=total - val, val_count=val,
                target_scale=targets.pop())


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _write_json(path, obj):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(obj, f, indent=1)
    os.replace(tmp, path)


def harvest(directory, seeds, n=128, h=0.2, patch=32, per_molecule=None, target_patches=200000, max_iterations=30,
            max_columns=8, patches_per_column=4, every=3, cg_tol=1e-6, cg_max_iter=400, functional=PZ81,
            base="kinetic"):
    """FP64 SCF of each molecule with the projected-CG harvest; one shard per
    seed.  Seeds whose shard exists are skipped, so an interrupted harvest
    resumes; resuming with a different grid, spacing, patch edge, functional
    or base is refused.  Returns the manifest."""
    os.makedirs(directory, exist_ok=True)
    per = per_molecule or max(1, -(-target_patches // max(len(seeds), 1)))
    target_scale = BASES[base][0]
    grid = Grid.centered(n, h)
    mpath = os.path.join(directory, "harvest.json")
    manifest = dict(format="dft-cnn-harvest", version=1, grid=n, h=h, patch=patch, per_molecule=per,
                    functional="PZ81" if functional == PZ81 else "PBE96", base=base, molecules=[])
    if os.path.exists(mpath):
        with open(mpath) as f:
            old = json.load(f)
        for key in ("grid", "h", "patch", "functional", "base"):
            if old.get(key, "diagonal" if key == "base" else None) != manifest[key]:
                raise ValueError(f"{mpath} was harvested with {key} = {old.get(key)!r}, not {manifest[key]!r}; "
                                 "use another directory")
        manifest["molecules"] = old.get("molecules", [])
    done = {m["seed"] for m in manifest["molecules"]}
    for seed in seeds:
        path = shard_path(directory, seed)
        if seed in done and os.path.exists(path):
            continue
        atoms = np.array(training_molecule(seed), dtype=float)
        ne = int(round(atoms[:, 0].sum()))
        t0 = time.perf_counter()
        tr = CnnTrainer(patch=patch, capacity=per, seed=seed)
        try:
            added = tr.harvest(grid, atoms, functional=functional, nstates=ne // 2 + 4, max_iterations=max_iterations,
                               poisson_method=POISSON_DIRECT, init_guess=INIT_HYDROGENIC, max_columns=max_columns,
                               patches_per_column=patches_per_column, every=every, cg_tol=cg_tol,
                               cg_max_iter=cg_max_iter, seed=seed, target_scale=target_scale)
            tr.save_patches(path, target_scale=target_scale, h=h)
        finally:
            tr.close()
        dt = time.perf_counter() - t0
        manifest["molecules"] = [m for m in manifest["molecules"] if m["seed"] != seed]
        manifest["molecules"].append(dict(seed=seed, atoms=len(atoms), electrons=ne, patches=int(added), seconds=dt))
        _write_json(mpath, manifest)
        _log(f"[harvest] seed {seed}: {len(atoms)} atoms, {ne} electrons, {added} patches, {dt:.1f} s")
    return manifest


def train(directory, steps=20000, batch=16, lr=1e-3, chunk=500, seed=1):
    """Adam on every shard but the last (the held-out molecule of the
    validation loss); writes model.npz and train_log.json."""
    plan = shard_plan(directory)
    tr = CnnTrainer(patch=plan["patch"], capacity=plan["total"], seed=seed)
    log = dict(format="dft-cnn-train", version=1, steps=steps, batch=batch, lr=lr, patches=plan["total"],
               train_patches=plan["train_count"], validation_patches=plan["val_count"], curve=[])
    try:
        for p in plan["paths"]:
            tr.load_patches(p)
        _log(f"[train] {plan['total']} patches of edge {plan['patch']} from {len(plan['paths'])} shards "
             f"({plan['val_count']} held out)")
        t0 = time.perf_counter()
        for s0 in range(0, steps, chunk):
            k = min(chunk, steps - s0)
            loss = tr.train(k, batch=batch, lr=lr, train_count=plan["train_count"])
            val = tr.eval(plan["train_count"], plan["val_count"]) if plan["val_count"] else None
            row = dict(step=s0 + k, train_loss=float(np.mean(loss)), validation_loss=val,
                       seconds=time.perf_counter() - t0)
            log["curve"].append(row)
            _log(f"[train] step {row['step']}: loss {row['train_loss']:.4e}"
                 + (f", held-out {val:.4e}" if val is not None else ""))
        model = tr.model(base=CNN_BASE_KINETIC if plan["target_scale"] == 2 else CNN_BASE_NONE)
    finally:
        tr.close()
    model.save(os.path.join(directory, "model.npz"))
    _write_json(os.path.join(directory, "train_log.json"), log)
    return model, log


def evaluate(directory, seeds, n=128, h=0.2, functional=PZ81, energy_tol=1e-5, max_iterations=60, smearing=0.01):
    """FP64 SCF of each held-out molecule with the correction the model builds
    on (the exact kinetic one for base 1, the diagonal one for base 0) and
    with the learned one (diagonal fallback per iteration); Fermi-Dirac
    smearing and Pulay mixing, as the random molecules are radicals.
    Records H applications, Davidson iterations, wall time and the energy
    difference, and writes each run's result document.  Returns the summary
    (also in eval.json)."""
    model = CnnModel.load(os.path.join(directory, "model.npz"))
    base_pc = PRECOND_KINETIC_EXACT if model.base == CNN_BASE_KINETIC else PRECOND_DIAG
    grid = Grid.centered(n, h)
    rows = []
    for seed in seeds:
        atoms = np.array(training_molecule(seed), dtype=float)
        ne = int(round(atoms[:, 0].sum()))
        runs = {}
        for name, pc, cnn in (("base", base_pc, None), ("learned", PRECOND_CNN, model)):
            r = scf_run(grid, atoms, functional=functional, precision=SCF_FP64, nstates=ne // 2 + 6,
                        max_iterations=max_iterations, poisson_method=POISSON_DIRECT, stop_rule=STOP_DE,
                        energy_tol=energy_tol, init_guess=INIT_HYDROGENIC, precond=pc, cnn=cnn,
                        smearing=smearing, mixer=MIX_PULAY, mix_depth=10)
            with open(os.path.join(directory, f"eval_seed{seed:05d}_{name}.json"), "w") as f:
                f.write(scf_result_json(grid, atoms, r))
            runs[name] = r
        d, c = runs["base"], runs["learned"]
        row = dict(seed=seed, atoms=len(atoms), electrons=ne, converged=[bool(d["converged"]), bool(c["converged"])],
                   happly=[int(d["happly_columns"]), int(c["happly_columns"])],
                   davidson_iterations=[int(d["davidson_iterations"]), int(c["davidson_iterations"])],
                   seconds=[d["seconds_total"], c["seconds_total"]],
                   cnn_applied=int(c["cnn_applied"]), cnn_fallbacks=int(c["cnn_fallbacks"]),
                   energy_difference=c["total"] - d["total"],
                   happly_gain=d["happly_columns"] / max(c["happly_columns"], 1),
                   time_gain=d["seconds_total"] / c["seconds_total"] if c["seconds_total"] > 0 else None)
        rows.append(row)
        _log(f"[evaluate] seed {seed}: H applies {row['happly'][0]} -> {row['happly'][1]} "
             f"(gain {row['happly_gain']:.2f}x), time {row['seconds'][0]:.1f} -> {row['seconds'][1]:.1f} s, "
             f"dE {row['energy_difference']:.2e} Ha")
    summary = dict(format="dft-cnn-eval", version=1, grid=n, h=h, base="kinetic" if model.base else "diagonal",
                   molecules=rows,
                   mean_happly_gain=float(np.mean([r["happly_gain"] for r in rows])) if rows else None,
                   max_abs_energy_difference=float(max(abs(r["energy_difference"]) for r in rows)) if rows else None)
    _write_json(os.path.join(directory, "eval.json"), summary)
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m dft_amd.cnn_pipeline", description=__doc__.split("\n\n")[0])
    ap.add_argument("step", choices=["harvest", "train", "evaluate", "all"])
    ap.add_argument("--dir", required=True, help="directory of shards, model and reports")
    ap.add_argument("--seeds", default=TRAIN_SEEDS, help="training seeds (harvest), e.g. 1-64")
    ap.add_argument("--eval-seeds", default=EVAL_SEEDS, help="held-out seeds (evaluate)")
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--h", type=float, default=0.2)
    ap.add_argument("--patch", type=int, default=32)
    ap.add_argument("--target-patches", type=int, default=200000)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--base", default="kinetic", choices=sorted(BASES))
    ap.add_argument("--max-iterations", type=int, default=30, help="SCF steps of each harvest run")
    a = ap.parse_args(argv)
    if a.step in ("harvest", "all"):
        harvest(a.dir, parse_seeds(a.seeds), n=a.grid, h=a.h, patch=a.patch, target_patches=a.target_patches,
                base=a.base, max_iterations=a.max_iterations)
    if a.step in ("train", "all"):
        train(a.dir, steps=a.steps, batch=a.batch, lr=a.lr)
    if a.step in ("evaluate", "all"):
        evaluate(a.dir, parse_seeds(a.eval_seeds), n=a.grid, h=a.h)
    return 0


if __name__ == "__main__":
    sys.exit(main())
