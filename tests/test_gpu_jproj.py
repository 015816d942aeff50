"""GPU: J-projected TRL, main_sm with target_J = [2J] (doubleLanczos).

The lowest states of a given J are compared with the dense eigenvectors of the
brute-force H in the 2M = 2J sector, classified by their <J^2> (ladder-operator
J^2): ckpot Be8 (J = 1, 2) and usdb Ne20 (J = 0, 2, 4) and the odd nucleus
usdb Na21 (J = 3/2), to 1e-7 MeV (FP64) and 1e-5 MeV (mixed).  Every returned
state has that J (from its Ritz vector's <J^2>), and the start vector is
always projected (n_jproj >= 1).
"""
import numpy as np
import pytest

import brute_force as bf
import smgpu
from smgpu.shellmodel import j_from_jj1

pytestmark = pytest.mark.gpu


def _dense_levels(path, nuc, Z, N, twoJ, k):
    H, _ = bf.brute_hamiltonian(path, nuc, Z, N, twoM=twoJ)
    J2 = bf.brute_j2(path, nuc, Z, N, twoM=twoJ)
    ev, U = np.linalg.eigh(H)
    out = []
    for i in range(len(ev)):
        if j_from_jj1(float(U[:, i] @ J2 @ U[:, i]), 1e-6) == 0.5 * twoJ:
            out.append(ev[i])
            if len(out) == k:
                break
    return np.array(out)


@pytest.mark.parametrize("fname,nuc,Z,N,twoJ,k", [("ckpot.snt", "Be8", 4, 4, 2, 2), ("ckpot.snt", "Be8", 4, 4, 4, 3),
                                                 ("usdb.snt", "Ne20", 10, 10, 0, 3), ("usdb.snt", "Ne20", 10, 10, 4, 3),
                                                 ("usdb.snt", "Ne20", 10, 10, 8, 2), ("usdb.snt", "Na21", 11, 10, 3, 3)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-7), ("mixed", 1e-5)])
def test_j_projected_levels(golden, fname, nuc, Z, N, twoJ, k, precision, tol):
    path = f"{golden}/{fname}"
    ref = _dense_levels(path, nuc, Z, N, twoJ, k)
    en, J, st = smgpu.main_sm(path, nuc, k, target_J=[twoJ], precision=precision, return_J=True,
                              return_stats=True, ls=max(k, 10), lm=60)
    np.testing.assert_allclose(en, ref, atol=tol)
    assert all(x == 0.5 * twoJ for x in J), J
    assert st["n_jproj"] >= 1


def test_main_sm_verbose_lines(golden, capsys):
    """verbose: main_sm's printed lines (shellmodel_main.jl:256-266): the dimension
    line, J of every state, En. and Ex. with %10.4f, or 2J= and En. with %12.5f for
    a J-projected solve; the printed values are the returned ones."""
    en, J = smgpu.main_sm(f"{golden}/ckpot.snt", "Be8", 4, verbose=True, return_J=True)
    out = capsys.readouterr().out
    assert out.startswith("   Be8      Z,N=(4,4)     c(2,2)     v(2,2)  mdim:           51 (  1.71 ) J ["), repr(out)
    lines = out.split("\n")
    assert lines[1] == "En. " + "".join("%10.4f " % x for x in en)
    assert lines[2] == "Ex. " + "".join("%10.4f " % (x - en[0]) for x in en)
    assert J[0] == 0.0
    en2 = smgpu.main_sm(f"{golden}/ckpot.snt", "Be8", 2, [4], verbose=True)
    out = capsys.readouterr().out
    assert "2J =   4" in out and out.rstrip("\n").endswith("2J= 4  En." + "".join("%12.5f " % x for x in en2))
