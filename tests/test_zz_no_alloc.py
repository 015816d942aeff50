"""smgpu_lanczos_steps_noalloc against smgpu_lanczos_steps, through the C ABI.

The no-allocation entry point keeps every device buffer of the call in the
library's static scratch array and runs on the handle's stream; it must give the
same alpha / beta as the allocating entry point (the same kernels in the same
order, so bit-identical), on two calls in a row whose second is smaller (Mg24,
mdim 28,503, then Ne20, mdim 640: the scratch then holds the first call's
contents), in FP64 and in mixed precision (FP32 basis; Ne20 takes the
one-workgroup step).  A size whose buffers exceed the scratch is rejected with
SMGPU_EINVAL and a message, and the entry point still works afterwards.
"""
import numpy as np
import pytest

import smgpu
from smgpu import synth
from smgpu._lib import lib, ptr

SMGPU_EINVAL = -1   # include/smgpu.h

pytestmark = pytest.mark.gpu

CASES = [("sd", 4, 4, 1, (-4.0, 0.0), 1.5, 60), ("sd", 2, 2, 11, (-6.0, 0.0), 1.2, 30)]


def nuc_tag(space, zv, nv):
    return f"{space}{zv}{nv}"


def _steps(fn, h, prec, nsteps, v0):
    a = np.zeros(nsteps)
    b = np.zeros(nsteps)
    rc = fn(h._h, prec, nsteps, ptr(v0), ptr(a), ptr(b))
    return rc, a, b


def test_noalloc_equals_lanczos_steps(snt_dir):
    L = lib()
    assert L.smgpu_lanczos_scratch_bytes() == 64 << 20
    setups = []
    for space, zv, nv, soff, spe, sig, nsteps in CASES:
        path, nuc = synth.config_snt(f"{snt_dir}/noalloc_{nuc_tag(space, zv, nv)}.snt", space, zv, nv,
                                     synth.BASE_SEED + soff, spe, sig)
        p = smgpu.Problem(path, nuc)
        h = smgpu.Hamiltonian(p)
        v0 = np.random.default_rng(zv * 10 + nv).standard_normal(p.mdim)
        print("noalloc case", nuc, "mdim", p.mdim, "nsteps", nsteps, flush=True)
        setups.append((nuc, p, h, v0, nsteps))
    assert setups[1][1].mdim < setups[0][1].mdim
    for prec in (0, 1):
        ref = []
        for nuc, p, h, v0, nsteps in setups:
            rc, a, b = _steps(L.smgpu_lanczos_steps, h, prec, nsteps, v0)
            assert rc == 0, L.smgpu_last_error()
            ref.append((a, b))
        # two no-allocation calls in a row, the second on the smaller basis, then the
        # first again (the scratch holds the previous call's contents each time)
        for k in (0, 1, 0):
            nuc, p, h, v0, nsteps = setups[k]
            print("noalloc call", nuc, "precision", prec, flush=True)
            rc, a, b = _steps(L.smgpu_lanczos_steps_noalloc, h, prec, nsteps, v0)
            assert rc == 0, L.smgpu_last_error()
            assert np.array_equal(a, ref[k][0]) and np.array_equal(b, ref[k][1]), (nuc, prec)
            assert np.all(np.isfinite(a)) and np.all(b > 0)


def test_noalloc_rejects_oversize(snt_dir):
    L = lib()
    path, nuc = synth.config_snt(f"{snt_dir}/noalloc_big.snt", "sd", 4, 4, synth.BASE_SEED + 1, (-4.0, 0.0), 1.5)
    p = smgpu.Problem(path, nuc)
    h = smgpu.Hamiltonian(p)
    v0 = np.ones(p.mdim)
    # (400 + 8) basis-sized FP64 vectors of 28,544 amplitudes: 93 MB > 64 MiB
    rc, _, _ = _steps(L.smgpu_lanczos_steps_noalloc, h, 0, 400, v0)
    assert rc == SMGPU_EINVAL
    msg = L.smgpu_last_error().decode()
    assert "scratch" in msg and "maximum" in msg, msg
    rc, a, b = _steps(L.smgpu_lanczos_steps_noalloc, h, 0, 20, v0)
    assert rc == 0
    rc, a2, b2 = _steps(L.smgpu_lanczos_steps, h, 0, 20, v0)
    assert rc == 0 and np.array_equal(a, a2) and np.array_equal(b, b2)
