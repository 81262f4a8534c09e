"""The reference's iteration checkpoint (pyHALMA.py:1813-1821 writes it, :405-438 reads it):
an ``np.savez`` archive ``HALMA_checkpoint.npz`` of plain arrays.

  info                    [iteration, cosmo_time, number of haloes]
  omm, npart              per halo: mass and member count
  oripas_before           the haloes' particle IDs, concatenated in halo order
  part_ih_before, part_oripas_before, part_insitu_before       per particle
  part_ih_before2, part_oripas_before2, part_insitu_before2    two iterations back
  pro1_before2            per halo: main progenitor two iterations back

``save`` writes these keys and ``load`` reads them with ``allow_pickle=False`` (the
archive holds no objects) and rebuilds the per-halo ID lists as the reference's loader
does.  ``old_catalogue`` gives what ``engine.progenitors`` takes for the old haloes.
"""
import os

import numpy as np

KEYS = ("info", "omm", "npart", "oripas_before", "part_ih_before", "part_oripas_before",
        "part_insitu_before", "part_ih_before2", "part_oripas_before2",
        "part_insitu_before2", "pro1_before2")


def save(directory, iteration, cosmo_time, omm, npart, oripas_before, **per_particle):
    """Write ``directory/HALMA_checkpoint.npz`` (pyHALMA.py:1813-1821).  ``oripas_before``
    is the concatenated ID array or a list of per-halo arrays; the per-particle and
    two-iterations-back arrays are keyword arguments named as in the archive (missing
    ones are stored empty)."""
    if isinstance(oripas_before, (list, tuple)):
        oripas_before = (np.concatenate([np.asarray(a) for a in oripas_before])
                         if len(oripas_before) else np.zeros(0, np.int64))
    unknown = set(per_particle) - set(KEYS)
    if unknown:
        raise ValueError(f"unknown checkpoint keys: {sorted(unknown)}")
    arrays = {k: np.asarray(per_particle.get(k, np.array([]))) for k in KEYS[4:]}
    info = np.array([iteration, cosmo_time, len(np.asarray(npart))])
    path = os.path.join(directory, "HALMA_checkpoint")
    np.savez(path, info=info, omm=np.asarray(omm), npart=np.asarray(npart),
             oripas_before=np.asarray(oripas_before), **arrays)
    return path + ".npz"


def load(path):
    """Read a checkpoint archive (pyHALMA.py:405-438): a dict of its arrays plus
    ``iteration``, ``cosmo_time``, ``n_haloes`` and ``oripas_list`` (per-halo IDs)."""
    if os.path.isdir(path):
        path = os.path.join(path, "HALMA_checkpoint.npz")
    with np.load(path, allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    info = d["info"]
    d["iteration"] = int(info[0])
    d["cosmo_time"] = float(info[1])
    d["n_haloes"] = int(info[2])
    npart = d["npart"].astype(np.int32)
    bounds = np.concatenate([[0], np.cumsum(npart, dtype=np.int64)])
    d["oripas_list"] = [d["oripas_before"][bounds[i]:bounds[i + 1]]
                        for i in range(d["n_haloes"])]
    return d


def old_catalogue(ckpt):
    """(old_npart, old_oripas, omm) of a loaded checkpoint, for engine.progenitors."""
    n = ckpt["n_haloes"]
    return ckpt["npart"][:n].astype(np.int64), ckpt["oripas_before"], ckpt["omm"][:n]
