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
  when N is not numerically positive definite.  Here the vectors stay resident
  and H_ij costs one H.v per sample on the GPU.
* prepEC's "TD" mode (eigenvector_continuation.jl:212-312, calcOBTD :1035,
  prepTBTD :1101, calcTBTD :1306): `prep_ec_td` reads the sample .wav files and
  writes the overlap matrix and the transition densities of every sample pair,
  one per interaction parameter (proton SPEs, neutron SPEs, then the TBME lines
  of the file), in the reference's tdmat layout; `solve_ec_td` (solveEC,
  :347-483) then builds H_ij = sum_k ME_k TD_k(i, j) for any interaction on the
  same orbits without an H.v.  The Hamiltonian is linear in its parameters,
  H(p) = sum_k p_k O_k, so TD_k(i, j) = <psi_i|O_k|psi_j> =
  <psi_i|H(b + e_k) - H(b)|psi_j> for any base b; with a generic b (every
  parameter nonzero) all the H(b + e_k) share the pp/nn sparsity pattern of H(b)
  and run as variant handles on one device index, each applied to all samples
  in shared multi-vector H.v launches (FP64).
"""
from __future__ import annotations

import os
import re
import tempfile
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


# ----------------------------------------------------------------- TD mode
@dataclass
class SntParams:
    """The parameters of an interaction file as readsmsnt reads them
    (shellmodel_main.jl:293-365): SPEs (proton orbits, then neutron orbits) and
    the TBME lines as written; massop scaling is applied by scaled_tbmes."""
    lp: int
    ln: int
    head: list               # model-space line and orbit lines, verbatim
    spe: np.ndarray          # [lp + ln]
    labels: list             # (i, j, k, l, 2J... as written: J) per TBME line
    tbme: np.ndarray         # [ntbme] as written
    massop: int
    Aref: float
    pw: float

    @property
    def n_params(self):
        return self.lp + self.ln + len(self.tbme)


def _snt_lines(path):
    out = []
    for line in open(path):
        t = line.strip()
        if t and t[0] in "!#":
            continue
        out.append(t)
    return out


def read_snt_params(path) -> SntParams:
    L = _snt_lines(path)
    lp, ln, _, _ = (int(x) for x in L[0].split()[:4])
    o = 1 + lp + ln
    head = L[:o]
    spe = np.zeros(lp + ln)
    nsp = int(L[o].split()[0])
    for i in range(nsp):
        t = L[o + 1 + i].split()
        spe[int(t[0]) - 1] = float(t[2])        # proton orbits 1..lp, neutron lp+1..lp+ln
    t = L[o + 1 + nsp].split()
    ntb, massop, Aref = int(t[0]), int(t[1]), float(t[2])
    pw = float(t[3]) if len(t) > 3 else 0.0
    labels, tb = [], np.zeros(ntb)
    for k in range(ntb):
        r = L[o + 2 + nsp + k].split()
        labels.append(tuple(int(x) for x in r[:5]))
        tb[k] = float(r[5])
    return SntParams(lp, ln, head, spe, labels, tb, massop, Aref, pw)


def mass_number(target_nuc):
    m = re.search(r"\d+", target_nuc)
    if not m:
        raise ValueError(f"nucleus {target_nuc!r} has no mass number")
    return int(m.group())


def scaled_tbmes(p: SntParams, target_nuc, ignore_massop=False):
    """oTBMEs: the TBMEs with readsmsnt's mass scaling (A/Aref)^p when massop == 1."""
    if p.massop == 1 and not ignore_massop:
        return p.tbme * (mass_number(target_nuc) / p.Aref) ** p.pw
    return p.tbme.copy()


def write_snt_params(path, p: SntParams, spe, tbme):
    """An interaction file on p's orbits with these SPEs and TBMEs taken literally
    (massop 0)."""
    with open(path, "w") as f:
        f.write("\n".join(p.head) + "\n")
        f.write(f"{p.lp + p.ln} 0\n")
        for i, e in enumerate(spe):
            f.write(f"{i + 1} {i + 1} {float(e)!r}\n")
        f.write(f"{len(tbme)} 0 {p.Aref!r} {p.pw!r}\n")
        for lab, v in zip(p.labels, tbme):
            f.write(" ".join(str(x) for x in lab) + f" {float(v)!r}\n")
    return path


def transition_densities(vectors, sntf, target_nuc, twoM, *, device=0, seed=20260101):
    """Overlaps N[i, j] = <psi_i|psi_j> and transition densities TD[k, i, j] =
    <psi_i|O_k|psi_j> of the sample vectors (rows of `vectors`, 2M = twoM) for every
    parameter k of sntf's orbits and TBME lines: k < lp + ln the number operator of
    that orbit (calcOBTD), then the operator multiplying TBME line k (calcTBTD, the
    mass-scaled oTBME).  Computed as differences of H.v over a generic base."""
    V = np.ascontiguousarray(np.atleast_2d(vectors), dtype=np.float64)
    p = read_snt_params(sntf)
    P = p.n_params
    rng = np.random.default_rng(seed)
    base = rng.uniform(0.5, 1.5, P) * rng.choice([-1.0, 1.0], P)
    tmp = tempfile.mkdtemp(prefix="smgpu_td_")

    def file_for(vals, name):
        return write_snt_params(os.path.join(tmp, name), p, vals[:p.lp + p.ln], vals[p.lp + p.ln:])

    prob = Problem(file_for(base, "base.snt"), target_nuc, twoM=int(twoM))
    if prob.mdim != V.shape[1]:
        raise ValueError(f"sample vectors have {V.shape[1]} amplitudes, the basis {prob.mdim}")
    ham = Hamiltonian(prob, device)

    def overlaps(h):
        M = V @ h.operate_H_multi(V, "fp64").T
        return 0.5 * (M + M.T)

    M0 = overlaps(ham)
    TD = np.zeros((P, V.shape[0], V.shape[0]))
    unit = None
    for k in range(P):
        vals = base.copy()
        vals[k] += 1.0
        prob.set_interaction(file_for(vals, f"unit_{k}.snt"))
        try:
            if unit is None:
                unit = ham.variant(prob)
            else:
                unit.update_interaction(prob)
            h = unit
        except Exception:                 # a pattern other than the base's: own tables
            h = Hamiltonian(Problem(prob.sntf, target_nuc, twoM=int(twoM)), device)
        TD[k] = overlaps(h) - M0
        os.remove(prob.sntf)
    return V @ V.T, TD


def idx_from_ij(i, j, dim):
    """1-based pair index of i <= j (shellmodel_main.jl:1297-1299)."""
    return (dim + 1) * (i - 1) - (i * (i - 1)) // 2 + j - i + 1


def write_tdmat(path, Nmat, TD):
    """The reference's TD file (eigenvector_continuation.jl:276-305), native Int64 /
    Float64: Dim, number of pairs, lTD, then per pair i <= j (1-based): i, j, Nij,
    N[i, j], TD[:, i, j]."""
    dim = Nmat.shape[0]
    P = TD.shape[0]
    with open(path, "wb") as f:
        np.array([dim, dim * (dim + 1) // 2, P], dtype="<i8").tofile(f)
        for i in range(1, dim + 1):
            for j in range(i, dim + 1):
                np.array([i, j, idx_from_ij(i, j, dim)], dtype="<i8").tofile(f)
                np.array([Nmat[i - 1, j - 1]], dtype="<f8").tofile(f)
                np.ascontiguousarray(TD[:, i - 1, j - 1], dtype="<f8").tofile(f)
    return path


def read_tdmat(path):
    """read_tdmat (eigenvector_continuation.jl:43-61): (Dim, ln, lTD, Nmat, infos, TDs)."""
    with open(path, "rb") as f:
        dim, npair, lTD = (int(x) for x in np.fromfile(f, dtype="<i8", count=3))
        Nmat = np.zeros((dim, dim))
        infos, TDs = [], []
        for _ in range(npair):
            i, j, nij = (int(x) for x in np.fromfile(f, dtype="<i8", count=3))
            t = float(np.fromfile(f, dtype="<f8", count=1)[0])
            Nmat[i - 1, j - 1] = Nmat[j - 1, i - 1] = t
            infos.append((i, j, nij))
            TDs.append(np.fromfile(f, dtype="<f8", count=lTD))
    return dim, npair, lTD, Nmat, infos, TDs


def read_sample_wavs(mdim, inpfs, tJ, num):
    """readRvecs! (eigenvector_continuation.jl:83-99) over the files: the first `num`
    vectors of each whose 2J equals tJ, and (file, nth, TD index) per vector."""
    from .wav import read_wav
    vecs, infos = [], []
    for inpf in inpfs:
        _, jj, v, _ = read_wav(inpf, mdim)
        for i in range(min(num, len(jj))):
            if int(jj[i]) == int(tJ):
                vecs.append(v[i])
                infos.append((inpf, i + 1, len(vecs)))
    return np.array(vecs), infos


def prep_ec_td(Hs, target_nuc, num_ev, num_ECsample, tJ, *, path_to_samplewav="", tdmatdir="./tdmat/", device=0):
    """prepEC(Hs, target_nuc, num_ev, num_ECsample, tJ, "TD"): reads the sample files
    {path}{nuc}_tmp_{i}_j{tJ}.wav (i < num_ECsample, as written by prepEC's sample
    mode with save_wav), writes {tdmatdir}wfinfos_{nuc}_J{tJ}.dat and the TD file
    {tdmatdir}{nuc}_J{tJ}.dat for the orbits and TBME lines of Hs[0].  Returns the
    TD file's path."""
    os.makedirs(tdmatdir, exist_ok=True)
    prob = Problem(Hs[0], target_nuc, twoM=int(tJ))
    inpfs = [f"{path_to_samplewav}{target_nuc}_tmp_{i}_j{tJ}.wav" for i in range(num_ECsample)]
    V, infos = read_sample_wavs(prob.mdim, inpfs, tJ, num_ev)
    if V.shape[0] == 0:
        raise ValueError("no sample vector with the target J")
    del prob
    with open(os.path.join(tdmatdir, f"wfinfos_{target_nuc}_J{tJ}.dat"), "w") as f:
        for inpf, nth, k in infos:
            f.write(f"{os.path.basename(inpf)} nth= {nth} TDmatidx  {k}\n")
    Nmat, TD = transition_densities(V, Hs[0], target_nuc, tJ, device=device)
    return write_tdmat(os.path.join(tdmatdir, f"{target_nuc}_J{tJ}.dat"), Nmat, TD)


def ec_hmat(TDs, infos, dim, sntf, target_nuc, ignore_massop=False):
    """H_ij = TD(i, j) . [SPE_p, SPE_n, oTBMEs] (solveEC, eigenvector_continuation.jl:430-436)."""
    p = read_snt_params(sntf)
    MEs = np.concatenate([p.spe, scaled_tbmes(p, target_nuc, ignore_massop)])
    H = np.zeros((dim, dim))
    for (i, j, _), td in zip(infos, TDs):
        if len(td) != len(MEs):
            raise ValueError(f"{sntf}: {len(MEs)} parameters, the TD file has {len(td)}")
        H[i - 1, j - 1] = H[j - 1, i - 1] = float(np.dot(td, MEs))
    return H


def solve_ec_td(Hs, target_nuc, tJNs, *, tdmatpath="./"):
    """solveEC(Hs, target_nuc, tJNs; tdmatpath) from TD files: for each [tJ, n] and
    each interaction, the lowest n eigenvalues of H c = E N c with H_ij from the
    transition densities (no H.v).  Returns [(sntf, tJ, vals)]."""
    out = []
    for tJ, nev in tJNs:
        dim, _, _, Nmat, infos, TDs = read_tdmat(os.path.join(tdmatpath, f"{target_nuc}_J{tJ}.dat"))
        if dim + 1 <= nev:
            raise ValueError(f"sample number must be larger than {nev}")
        Linv = np.linalg.inv(_cholesky_shifted(Nmat))
        for sntf in Hs:
            H = ec_hmat(TDs, infos, dim, sntf, target_nuc)
            vals = np.linalg.eigvalsh(Linv @ H @ Linv.T)[:nev]
            out.append((sntf, tJ, vals))
    return out
