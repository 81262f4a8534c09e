"""Checkpoint archive (pyHALMA.py:1813-1821 / :405-438): round trip of the reference's
keys and per-halo ID lists, no pickled objects."""
import numpy as np
import pytest

from pyhalma_amd import checkpoint


def test_round_trip(tmp_path):
    ids = [np.array([5, 9, 2]), np.array([7]), np.array([11, 12])]
    path = checkpoint.save(str(tmp_path), 2550, 1.25, omm=np.array([3.0, 1.0, 2.0]),
                           npart=np.array([3, 1, 2]), oripas_before=ids,
                           part_ih_before=np.array([0, 1, 2]), pro1_before2=np.array([1, 0]))
    d = checkpoint.load(path)
    assert d["iteration"] == 2550 and d["cosmo_time"] == 1.25 and d["n_haloes"] == 3
    assert [a.tolist() for a in d["oripas_list"]] == [[5, 9, 2], [7], [11, 12]]
    assert set(checkpoint.KEYS) <= set(d)
    assert d["part_oripas_before"].size == 0
    npart, oripas, omm = checkpoint.old_catalogue(checkpoint.load(str(tmp_path)))
    assert npart.tolist() == [3, 1, 2] and oripas.tolist() == [5, 9, 2, 7, 11, 12]
    assert omm.tolist() == [3.0, 1.0, 2.0]
    with pytest.raises(ValueError):
        checkpoint.save(str(tmp_path), 1, 0.0, [], [], [], bogus=np.zeros(1))
