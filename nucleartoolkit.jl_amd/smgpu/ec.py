"""Eigenvector continuation over the GPU solver (one basis, many interactions).

Mirrors prepEC's "sample" mode and solveEC
(src/ShellModel/eigenvector_continuation.jl:102-212, 347-483):

* `prep_ec(Hs, target_nuc, num_ev, tJ)` solves the lowest states of every
  sample interaction on one M-scheme basis (2M = tJ, as prepEC's Mtot = tJ).
  The index tables are built and uploaded once, and only the coefficients are
  swapped per interaction (smgpu_problem_set_interaction /
  smgpu_ham_update_interaction).  With tJ given, every solve is J-projected
  Lanczos (smgpu_trl_solve_j, as prepEC's TRL with doubleLanczos), so the
  num_ev lowest states of J = tJ/2 are the EC basis; without it, the lowest
  num_ev states of 2M = A % 2 are.  Without tJ the interactions are solved in
  lockstep groups of `batch` (variant handles on the one index,
  smgpu_trl_solve_variants: the H.v of a group shares launches), each solve
  identical to the one-at-a-time solve with the same seed.  Optionally each
  sample is written as a .wav file like prepEC's save_wav.
* `solve_ec(Hs, target_nuc, samples, num_ev_target)` builds, for every target
  interaction, H_ij = <psi_i|H|psi_j> and N_ij = <psi_i|psi_j> and solves
  H c = E N c through the Cholesky factor of N, with solveEC's epsilon shift
  when N is not numerically positive definite.  The reference evaluates H_ij
  from transition densities (prepEC "TD" mode) so that the sample vectors need
  not be kept.  Here the vectors stay resident and H_ij costs one H.v per
  sample on the GPU.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from .shellmodel import Hamiltonian, Problem, j_from_jj1, trl_variants
from .wav import write_ritzvecs


@dataclass
class ECSamples:
    nucleus: str
    twoM: int
    vectors: np.ndarray                  # [n_samples, mdim]
    energies: list = field(default_factory=list)   # per interaction: lowest num_ev energies
    Js: list = field(default_factory=list)         # per interaction: J of each state
    source: list = field(default_factory=list)     # (interaction file, state index) per sample


def prep_ec(Hs, target_nuc, num_ev, tJ=None, *, lm=300, ls=20, tol=1e-8, precision="fp64", device=0,
            seed=1234, save_wav=False, path_to_samplewav="", batch=4):
    """prepEC(Hs, target_nuc, num_ev, num_ECsample, tJ, "sample"): sample vectors
    of every interaction in Hs on the basis of 2M = tJ (default A % 2)."""
    prob = Problem(Hs[0], target_nuc, twoM=-1 if tJ is None else int(tJ))
    twoM = int(prob.info.twoM)
    ham = Hamiltonian(prob, device)
    j2 = prob.j2_hamiltonian(device) if tJ is not None else None
    nev = min(int(num_ev), prob.mdim)
    # (energies, Ritz vectors) per interaction, in the order of Hs
    solved = []
    if tJ is None and batch > 1:
        for g in range(0, len(Hs), batch):
            group = []
            for k in range(g, min(g + batch, len(Hs))):
                if k == 0:
                    group.append(ham)
                else:
                    prob.set_interaction(Hs[k])
                    group.append(ham.variant(prob))
            out = trl_variants(group, nev, [seed] * len(group), lm=lm, ls=ls, tol=tol, precision=precision,
                               want_vectors=True)
            solved.extend((ev, R) for ev, R, _ in out)
    else:
        for k, sntf in enumerate(Hs):
            if k > 0:
                prob.set_interaction(sntf)
                ham.update_interaction(prob)
            ev, R, _ = ham.trl(nev, lm=lm, ls=ls, tol=tol, precision=precision, seed=seed, want_vectors=True,
                               j2=j2, twoJ=None if tJ is None else int(tJ))
            solved.append((ev, R))
    vecs, src, energies, Js = [], [], [], []
    hj = j2 if j2 is not None else prob.j2_hamiltonian(device)     # J^2 depends on the basis only
    for sntf, (ev, R) in zip(Hs, solved):
        J = np.array([j_from_jj1(x) for x in np.einsum("ij,ij->i", R, hj.operate_H_multi(R, "fp64"))])
        energies.append(ev)
        Js.append(J)
        for i in range(nev):
            if tJ is None or J[i] == 0.5 * tJ:
                vecs.append(R[i])
                src.append((sntf, i))
        if save_wav:
            csnt = os.path.basename(sntf).split(".")[0]
            oupf = f"{path_to_samplewav}{target_nuc}_{csnt}.wav"
            if tJ is not None:
                oupf = f"{path_to_samplewav}{target_nuc}_{csnt}_j{tJ}.wav"
            write_ritzvecs(oupf, twoM, ev, J, R)
    if not vecs:
        raise ValueError("no sample state has the requested J")
    return ECSamples(target_nuc, twoM, np.array(vecs), energies, Js, src)


def _cholesky_shifted(N):
    """myCholesky! with solveEC's epsilon prescription (eigenvector_continuation.jl:400-410)."""
    try:
        return np.linalg.cholesky(N)
    except np.linalg.LinAlgError:
        dim = N.shape[0]
        u = 1.1e-16
        shift = np.trace(N) * u * ((dim + 2) / (1.0 - ((dim + 1) * (dim + 2)) * u))
        return np.linalg.cholesky(N + shift * np.eye(dim))


def solve_ec(Hs, target_nuc, samples: ECSamples, num_ev_target, *, device=0, precision="fp64"):
    """solveEC: for each interaction in Hs, the lowest num_ev_target eigenvalues of
    H in the span of the sample vectors.  Returns [(sntf, 2J or 2M, vals)]."""
    V = samples.vectors
    dim = V.shape[0]
    if dim + 1 <= num_ev_target:
        raise ValueError(f"sample number must be larger than {num_ev_target}")
    Nmat = V @ V.T
    L = _cholesky_shifted(Nmat)
    Linv = np.linalg.inv(L)
    prob = Problem(Hs[0], target_nuc, twoM=samples.twoM)
    ham = Hamiltonian(prob, device)
    out = []
    for k, sntf in enumerate(Hs):
        if k > 0:
            prob.set_interaction(sntf)
            ham.update_interaction(prob)
        HV = ham.operate_H_multi(V, precision)
        Hmat = V @ HV.T
        Hmat = 0.5 * (Hmat + Hmat.T)
        tildH = Linv @ Hmat @ Linv.T
        vals = np.linalg.eigvalsh(tildH)[:num_ev_target]
        out.append((sntf, samples.twoM, vals))
    return out
