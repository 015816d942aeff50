"""GPU: Lanczos steps captured as a hipGraph and replayed (SMGPU_STEPS_GRAPH) give
the same solve as launched steps (SMGPU_STEPS_LAUNCH): the same eigenvalues and
Ritz vectors bit for bit, the same iteration and restart counts, and at most one
more (speculative) H·v.  Covered: the multi-kernel
step with partial reorthogonalisation (mixed, Mg24, mdim 28,503), the full local
reorthogonalisation (FP64), the one-workgroup step of small bases (mixed, mdim
640), thick restarts (lm 40 forces several), a solve whose Krylov space is
exhausted (mdim 51, every level), and a second solve on the same handle (the
captured step state reset by the next solve)."""
import numpy as np
import pytest

import smgpu
from smgpu import synth

pytestmark = pytest.mark.gpu


def _pair(h, **kw):
    a = h.trl(want_vectors=True, steps="launch", **kw)
    b = h.trl(want_vectors=True, steps="graph", **kw)
    return a, b


def _same(a, b):
    (ea, va, sa), (eb, vb, sb) = a, b
    assert np.array_equal(ea, eb), (ea - eb)
    assert np.array_equal(va, vb)
    for k in ("iterations", "n_restart"):
        assert sa[k] == sb[k], (k, sa[k], sb[k])
    # replayed steps test convergence one step later: a converging solve runs one
    # more speculative step (entries past the converged ones only)
    for k in ("n_hv", "n_reorth"):
        assert 0 <= sb[k] - sa[k] <= 1, (k, sa[k], sb[k])


@pytest.mark.parametrize("space,zv,nv,soff,nev,lm", [("sd", 4, 4, 1, 5, 300), ("sd", 4, 4, 1, 5, 40),
                                                   ("sd", 2, 2, 11, 4, 300), ("sd", 2, 2, 11, 4, 30)])
def test_graph_steps_bit_identical(snt_dir, space, zv, nv, soff, nev, lm):
    path, nuc = synth.config_snt(f"{snt_dir}/gs_{zv}{nv}.snt", space, zv, nv, synth.BASE_SEED + soff, (-4.0, 0.0), 1.5)
    p = smgpu.Problem(path, nuc)
    h = smgpu.Hamiltonian(p)
    for prec in ("mixed", "fp64"):
        a, b = _pair(h, num_ev=nev, lm=lm, ls=20, precision=prec)
        print(nuc, p.mdim, prec, "lm", lm, "iterations", a[2]["iterations"], "restarts", a[2]["n_restart"], flush=True)
        _same(a, b)
        if lm == 40:
            assert a[2]["n_restart"] > 0


def test_graph_steps_exhausted_space(golden):
    p = smgpu.Problem(f"{golden}/ckpot.snt", "Be8")
    h = smgpu.Hamiltonian(p)
    for prec in ("mixed", "fp64"):
        a, b = _pair(h, num_ev=p.mdim if p.mdim <= 32 else 10, lm=300, ls=20, precision=prec)
        _same(a, b)
    # two graph solves in a row on one handle, another seed: equal to launched steps
    a = h.trl(10, seed=99, steps="launch")
    b = h.trl(10, seed=99, steps="graph")
    assert np.array_equal(a[0], b[0])
