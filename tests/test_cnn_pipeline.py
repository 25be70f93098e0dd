"""Host logic of the learned-preconditioner pipeline (dft_amd/cnn_pipeline.py)
and the molecule set it draws from (dft_amd/geometry.py): seed lists, the
geometry constraints of every training and held-out molecule, and the shard
plan that splits training from validation.  The harvest, training and
evaluation steps themselves run on the GPU."""
import itertools

import numpy as np
import pytest

import dft_amd
from dft_amd import cnn_pipeline as cp
from dft_amd.geometry import random_cluster, training_molecule


def test_parse_seeds():
    assert cp.parse_seeds("1-64") == list(range(1, 65))
    assert cp.parse_seeds("1001-1008") == list(range(1001, 1009))
    assert cp.parse_seeds("3, 1-4,9") == [3, 1, 2, 4, 9]
    assert cp.parse_seeds("") == []
    for bad in ("5-2", "-3", "a-b"):
        with pytest.raises(ValueError):
            cp.parse_seeds(bad)


@pytest.mark.parametrize("seed", list(range(1, 65)) + list(range(1001, 1009)))
def test_training_molecules_meet_the_set_definition(seed):
    m = np.array(training_molecule(seed))
    z, pos = m[:, 0], m[:, 1:]
    assert set(z) <= {1.0, 6.0, 7.0, 8.0}
    # 4..12 drawn atoms, plus one H when their electron count is odd
    assert 4 <= len(m) <= 13 and int(z.sum()) % 2 == 0
    if len(m) == 13:
        assert z[-1] == 1.0
    assert np.all(np.linalg.norm(pos, axis=1) <= 5.0)
    d = [np.linalg.norm(a - b) for a, b in itertools.combinations(pos, 2)]
    assert min(d) >= 2.0
    assert np.array_equal(m, np.array(training_molecule(seed)))  # a seed names one geometry


def test_training_set_varies():
    sizes = {len(training_molecule(s)) for s in range(1, 65)}
    assert len(sizes) >= 5
    assert not np.array_equal(np.array(training_molecule(1)), np.array(training_molecule(2)))


def test_cluster_generator_keeps_the_benchmark_cluster():
    """The 40-atom cluster of the 256^3 workload: seed 5, radius 7, 150-170 electrons."""
    m = np.array(random_cluster(5, 40, 7.0, ne_range=(150, 170)))
    assert m.shape == (40, 4) and int(m[:, 0].sum()) == 168


def _shard(path, count, P):
    rng = np.random.default_rng(count)
    dft_amd.save_patch_shard(path, rng.standard_normal((count, 2, P ** 3)).astype(np.float32),
                             rng.standard_normal((count, P ** 3)).astype(np.float32))


def test_shard_plan(tmp_path):
    with pytest.raises(FileNotFoundError):
        cp.shard_plan(tmp_path)
    for seed, count in ((3, 5), (1, 7), (12, 4)):
        _shard(cp.shard_path(tmp_path, seed), count, 4)
    plan = cp.shard_plan(tmp_path)
    assert [p.rsplit("/", 1)[1] for p in plan["paths"]] == ["seed00001.dftp", "seed00003.dftp", "seed00012.dftp"]
    assert plan["counts"] == [7, 5, 4] and plan["patch"] == 4
    assert (plan["total"], plan["train_count"], plan["val_count"]) == (16, 12, 4)
    _shard(cp.shard_path(tmp_path, 20), 2, 5)
    with pytest.raises(ValueError, match="patch edges"):
        cp.shard_plan(tmp_path)


def test_single_shard_trains_on_everything(tmp_path):
    _shard(cp.shard_path(tmp_path, 1), 6, 4)
    plan = cp.shard_plan(tmp_path)
    assert (plan["train_count"], plan["val_count"]) == (6, 0)
