"""GPU: TRL solves in lockstep on one Hamiltonian (smgpu_trl_solve_batch).

Each round issues the H·v of every unfinished member in shared launches (up to
four vectors per launch), then every member runs its own Lanczos step.  Each
member's eigenvalues, Ritz vectors and iteration counts must equal, bit for bit,
the same solve run alone (the shared launches give each vector the one-vector
kernel's result, and the rest of a step is the member's own): sd and pf bases
(two-row and one-row kernel shapes), FP64 and mixed, 1..5 members, and members
that restart at different steps (lm 40) or converge at different steps.

Mixed batches take the projection passes of partial re-orthogonalisation only on
the steps whose omega gate runs them (the host reads the gate), while a solve
alone queues them gated on the device: the pass counts (n_reorth, and the DGKS
second passes n_reorth2) must agree too, with passes on some steps and not on all.
"""
import numpy as np
import pytest

import smgpu
from smgpu import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("space,zv,nv,seed", [("sd", 4, 4, 21), ("pf", 3, 3, 22)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_equals_solo(snt_dir, space, zv, nv, seed, precision):
    path, nuc = synth.config_snt(f"{snt_dir}/batch_{space}.snt", space, zv, nv, synth.BASE_SEED + seed, (-6, 0), 1.2)
    h = smgpu.Hamiltonian(smgpu.Problem(path, nuc))
    for seeds, kw in [([7], {}), ([1, 2], {}), ([3, 4, 5, 6, 8], {}), ([11, 12, 13], dict(lm=40, ls=10))]:
        out = h.trl_batch(5, seeds, precision=precision, want_vectors=True, **kw)
        for sd, (ev, vec, st) in zip(seeds, out):
            ev1, vec1, st1 = h.trl(5, precision=precision, seed=sd, want_vectors=True, **kw)
            assert np.array_equal(ev, ev1), (sd, ev - ev1)
            assert np.array_equal(vec, vec1)
            assert (st["iterations"], st["n_hv"], st["n_restart"]) == (st1["iterations"], st1["n_hv"],
                                                                        st1["n_restart"])
            assert (st["n_reorth"], st["n_reorth2"]) == (st1["n_reorth"], st1["n_reorth2"])
            if precision == "mixed":
                assert 0 < st["n_reorth"] < st["n_hv"], st
        if "lm" in kw:
            assert all(o[2]["n_restart"] > 0 for o in out)
