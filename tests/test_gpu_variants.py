"""GPU: interaction variants on one basis sharing one device index (prepEC's
samples, eigenvector_continuation.jl:140-175: one basis, an interaction per
sample).

* A variant handle (Hamiltonian.variant) applies the same H, bit for bit, as a
  handle uploaded from scratch for that interaction: the shared index carries
  nothing of the base interaction.
* smgpu_apply_H_host_variants: every row bit-identical to the variant's own
  smgpu_apply_H, FP64 and FP32, for orders that mix repeated and distinct
  handles (a run of one handle goes to the shared-coefficient kernels, the rest
  pair up in the per-vector-coefficient kernels, an odd one out goes alone), on
  sd Mg24 and pf Cr48 (different neutron block widths).
* smgpu_trl_solve_variants: each member's eigenvalues, Ritz vectors and step
  counts bit-identical to the same solve on its own handle alone, mixed and
  FP64, with a repeated handle among the members and members that restart.
* The solved levels of each variant agree with a dense diagonalisation of the
  variant's H on the sd basis of Ne20 (brute force, 1e-7 MeV).
"""
import numpy as np
import pytest

import brute_force as bf
import smgpu
from smgpu import synth

pytestmark = pytest.mark.gpu


def _variants(tmp, n, space, zv, nv):
    paths, nuc = synth.variant_snts(tmp, n, space=space, zv=zv, nv=nv)
    base = smgpu.Hamiltonian(smgpu.Problem(paths[0], nuc))
    hams = [base] + [base.variant(smgpu.Problem(p, nuc)) for p in paths[1:]]
    return paths, nuc, hams


@pytest.fixture(scope="module")
def mg24(tmp_path_factory):
    return _variants(str(tmp_path_factory.mktemp("var_sd")), 3, "sd", 4, 4)


@pytest.fixture(scope="module")
def cr48(tmp_path_factory):
    return _variants(str(tmp_path_factory.mktemp("var_pf")), 3, "pf", 4, 4)


def test_variant_equals_fresh_upload(mg24):
    paths, nuc, hams = mg24
    rng = np.random.default_rng(5)
    v = rng.standard_normal(hams[0].mdim)
    for p, h in zip(paths[1:], hams[1:]):
        fresh = smgpu.Hamiltonian(smgpu.Problem(p, nuc))
        for prec in ("fp64", "mixed"):
            assert np.array_equal(h.operate_H(v, prec), fresh.operate_H(v, prec))
    # the variants differ from the base (the test above is not vacuous)
    assert not np.array_equal(hams[1].operate_H(v), hams[0].operate_H(v))


@pytest.mark.parametrize("case", ["mg24", "cr48"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hv_variants_bit_identical(request, case, precision):
    _, _, hams = request.getfixturevalue(case)
    A, B, C = hams
    rng = np.random.default_rng(11)
    for order in ([A, B], [A, B, C], [A, A, B, C, C], [B, A, B, A], [C, A, A, A, A, B]):
        V = rng.standard_normal((len(order), A.mdim))
        W = smgpu.operate_H_variants(order, V, precision)
        for b, h in enumerate(order):
            assert np.array_equal(W[b], h.operate_H(V[b], precision)), (case, precision, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_trl_variants_equal_solo(mg24, precision):
    _, _, hams = mg24
    A, B, C = hams
    for members, seeds, kw in [([A, B, C], [1, 2, 3], {}), ([B, A, B, C], [4, 5, 6, 7], dict(lm=40, ls=10))]:
        out = smgpu.trl_variants(members, 5, seeds, precision=precision, want_vectors=True, **kw)
        for h, sd, (ev, vec, st) in zip(members, seeds, out):
            ev1, vec1, st1 = h.trl(5, precision=precision, seed=sd, want_vectors=True, **kw)
            assert np.array_equal(ev, ev1), (sd, ev - ev1)
            assert np.array_equal(vec, vec1)
            assert (st["iterations"], st["n_hv"], st["n_restart"]) == (st1["iterations"], st1["n_hv"],
                                                                        st1["n_restart"])
        if "lm" in kw:
            assert all(o[2]["n_restart"] > 0 for o in out)


def test_trl_variants_dense_levels(tmp_path):
    paths, nuc = synth.variant_snts(str(tmp_path), 3, space="sd", zv=2, nv=2)
    base = smgpu.Hamiltonian(smgpu.Problem(paths[0], nuc))
    hams = [base] + [base.variant(smgpu.Problem(p, nuc)) for p in paths[1:]]
    out = smgpu.trl_variants(hams, 6, [1, 2, 3], precision="mixed")
    for p, (ev, _, _) in zip(paths, out):
        H, _ = bf.brute_hamiltonian(p, nuc, 10, 10)
        np.testing.assert_allclose(ev, np.linalg.eigvalsh(H)[:6], atol=1e-5)
    out64 = smgpu.trl_variants(hams, 6, [1, 2, 3], precision="fp64")
    for p, (ev, _, _) in zip(paths, out64):
        H, _ = bf.brute_hamiltonian(p, nuc, 10, 10)
        np.testing.assert_allclose(ev, np.linalg.eigvalsh(H)[:6], atol=1e-7)


def test_variant_errors(mg24, tmp_path):
    """A variant needs the base's basis; variant launches and solves need handles on one
    index.  Both fail with EINVAL, and the handles stay usable."""
    paths, nuc, hams = mg24
    other_path, other_nuc = synth.config_snt(str(tmp_path / "o.snt"), "sd", 3, 3, synth.BASE_SEED + 93, (-6, 0),
                                             1.2)
    other = smgpu.Hamiltonian(smgpu.Problem(other_path, other_nuc))
    with pytest.raises(smgpu.SmgpuError):
        hams[0].variant(smgpu.Problem(other_path, other_nuc))
    fresh = smgpu.Hamiltonian(smgpu.Problem(paths[1], nuc))          # same basis, its own index
    V = np.ones((2, hams[0].mdim))
    with pytest.raises(smgpu.SmgpuError):
        smgpu.operate_H_variants([hams[0], fresh], V)
    with pytest.raises(smgpu.SmgpuError):
        smgpu.trl_variants([hams[0], fresh], 2, [1, 2])
    assert other.mdim != hams[0].mdim
    w = smgpu.operate_H_variants([hams[0], hams[1]], V)
    assert np.array_equal(w[1], hams[1].operate_H(V[1]))


def test_c5_cr48_variants_lockstep_vs_fp64(tmp_path):
    """The interaction-variant sweep at its own shape, on 8 of its variants: Cr48
    (pf 4p4n, mdim 1,963,461), base TBMEs N(0, 1.0) plus N(0, 0.05) on every TBME per
    variant, lowest 3 each, all 8 solved in one lockstep batch on one shared index with
    the multi-vector FP32 H·V (smgpu_trl_solve_variants).  Every eigenvalue of every
    variant agrees with the plain FP64 solve of that variant alone within 1e-5 MeV, and
    the variants' spectra differ from each other (the batch is not one interaction)."""
    paths, nuc = synth.variant_snts(str(tmp_path), 8)
    base = smgpu.Hamiltonian(smgpu.Problem(paths[0], nuc))
    assert base.mdim == 1963461
    hams = [base] + [base.variant(smgpu.Problem(p, nuc)) for p in paths[1:]]
    out = smgpu.trl_variants(hams, 3, list(range(100, 108)), precision="mixed")
    grounds = []
    for i, (h, (ev, _, st)) in enumerate(zip(hams, out)):
        ev64, _, _ = h.trl(3, precision="fp64", seed=200 + i)
        print("variant", i, ev, np.abs(ev - ev64).max(), st["iterations"], flush=True)
        assert len(ev) == 3 and np.abs(ev - ev64).max() <= 1e-5
        grounds.append(ev[0])
    assert np.ptp(grounds) > 1e-3
