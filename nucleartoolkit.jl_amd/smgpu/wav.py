"""Ritz-vector files in the reference's .wav layout.

writeRitzvecs (src/ShellModel/shellmodel_main.jl:1215-1229) writes, in the
host's native (little-endian) byte order:
    Int32 num_ev, Int32 Mtot (2M), Float64 energies[num_ev], Int32 2J[num_ev],
    then num_ev vectors of mdim Float64 in the M-scheme order.
read_wav / read_appwav (:979-1003) read the same layout; the latter takes the
leading vector(s) as a Lanczos start (main_sm's in_wf, :194-198).
"""
from __future__ import annotations

import numpy as np


def write_ritzvecs(path, mtot, energies, totJs, vecs):
    """writeRitzvecs: energies[num_ev], J values (written as Int32(2J)), vecs[num_ev, mdim]."""
    energies = np.asarray(energies, dtype=np.float64)
    vecs = np.ascontiguousarray(vecs, dtype=np.float64)
    n = len(energies)
    if vecs.shape[0] != n:
        raise ValueError("one vector per energy")
    js = np.array([int(round(2 * j)) for j in totJs], dtype=np.int32)   # Int32(2*totJs); J is a multiple of 1/2 or -1
    with open(path, "wb") as f:
        np.array([n, int(mtot)], dtype="<i4").tofile(f)
        energies.astype("<f8").tofile(f)
        js.astype("<i4").tofile(f)
        vecs.astype("<f8").tofile(f)
    return path


def read_wav(path, mdim):
    """read_wav(all=true): (energies, 2J values, vectors[num_ev, mdim], Mtot)."""
    with open(path, "rb") as f:
        n, mtot = np.fromfile(f, dtype="<i4", count=2)
        es = np.fromfile(f, dtype="<f8", count=n)
        jj = np.fromfile(f, dtype="<i4", count=n)
        v = np.fromfile(f, dtype="<f8", count=n * mdim)
    if v.size != n * mdim:
        raise ValueError(f"{path}: expected {n} vectors of length {mdim}")
    return es, jj, v.reshape(n, mdim), int(mtot)


def read_appwav(path, mdim, q=1):
    """read_appwav: the first q vectors of a .wav file (the Lanczos start)."""
    es, jj, v, mtot = read_wav(path, mdim)
    if len(es) < q:
        raise ValueError("#(appwav) must be >= q")
    return v[:q].copy() if q > 1 else v[0].copy()
