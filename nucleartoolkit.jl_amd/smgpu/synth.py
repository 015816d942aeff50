"""Seeded synthetic isospin-conserving interactions written as `.snt` files.

The writer follows the reference's own isospin -> proton/neutron conventions so
that the file it produces is an ordinary interaction file for `main_sm`:

* isospin labels are enumerated like `make_int`
  (src/ShellModel/input_int_snt.jl:8-61): kets (a<=b) with J in
  |ja-jb|/2..(ja+jb)/2, identical orbits keep even J for T=1 and odd J for
  T=0, and every upper-triangle (bra-index <= ket-index) pair with equal J
  becomes one label;
* pn lines get the factors and swap phases of `int2snt`
  (src/ShellModel/input_int_snt.jl:63-145);
* the header layout is the one `write_vs_snt` emits
  (src/IMSRG.jl/valencespace.jl:349-350): `ntbme 0 A` means no mass scaling.

Random numbers come from `numpy.random.Generator(PCG64(seed))`, drawn in this
order: the orbit SPEs (uniform, one per orbit, shared by protons and neutrons),
then every T=1 TBME, then every T=0 TBME (normal, in label order).
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

# (n, l, 2j) in the listed order; the first orbit is the first m-state block.
SPACES = {
    "p": [(0, 1, 1), (0, 1, 3)],
    "sd": [(0, 2, 5), (1, 0, 1), (0, 2, 3)],
    "pf": [(0, 3, 7), (1, 1, 3), (0, 3, 5), (1, 1, 1)],
}
CORES = {"p": (2, 2), "sd": (8, 8), "pf": (20, 20)}

ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si",
            "P", "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni",
            "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb",
            "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn"]


def nucleus_name(space: str, zv: int, nv: int) -> str:
    """Nucleus string ("Cr48") for `zv` valence protons and `nv` valence neutrons."""
    cp, cn = CORES[space]
    z = cp + zv
    return f"{ELEMENTS[z - 1]}{z + cn + nv}"


def isospin_labels(orbits):
    """T=1 and T=0 (a,b,c,d,J) labels, 1-based orbit indices, in make_int order."""
    def pair_list(identical_parity):
        out = []
        for i, (_, _, j1) in enumerate(orbits, start=1):
            for k, (_, _, j2) in enumerate(orbits, start=1):
                if k < i:
                    continue
                for J in range(abs(j1 - j2) // 2, (j1 + j2) // 2 + 1):
                    if i == k and J % 2 == identical_parity:
                        continue
                    out.append((i, k, J))
        return out

    def upper(pairs):
        labs = []
        for x in range(len(pairs)):
            for y in range(x, len(pairs)):
                if pairs[x][2] != pairs[y][2]:
                    continue
                labs.append((pairs[x][0], pairs[x][1], pairs[y][0], pairs[y][1], pairs[y][2]))
        return labs

    return upper(pair_list(1)), upper(pair_list(0))


def _pn_factor(orbits, line, labels, T):
    """Coefficient of the isospin TBME `labels[idx]` in pn line `line` (int2snt)."""
    a, b, c, d, J = line
    jj = [o[2] for o in orbits]
    for idx, (ta, tb, tc, td, tJ) in enumerate(labels):
        if tJ != J:
            continue
        dab = 1.0 if ta == tb else 0.0
        dcd = 1.0 if tc == td else 0.0
        sgn = (-1) ** tJ
        if T == 1:
            coef = 0.5 * math.sqrt((1 + dab * sgn) * (1 + dcd * sgn))
            extra = 1
        else:
            coef = 0.5 * math.sqrt((1 - dab * sgn) * (1 - dcd * sgn))
            extra = 0
        if (a, b, c, d) in ((ta, tb, tc, td), (tc, td, ta, tb)):
            return idx, coef
        if (a, b, c, d) in ((tb, ta, tc, td), (tc, td, tb, ta)):
            ja, jb = jj[tb - 1], jj[ta - 1]
            return idx, (-1) ** ((ja + jb) // 2 + tJ + extra) * coef
        if (a, b, c, d) in ((ta, tb, td, tc), (td, tc, ta, tb)):
            jc, jd = jj[td - 1], jj[tc - 1]
            return idx, (-1) ** ((jc + jd) // 2 + tJ + extra) * coef
        if (a, b, c, d) in ((tb, ta, td, tc), (td, tc, tb, ta)):
            ja, jb = jj[tb - 1], jj[ta - 1]
            jc, jd = jj[td - 1], jj[tc - 1]
            return idx, (-1) ** ((ja + jb + jc + jd) // 2) * coef
    return -1, 0.0


@dataclass
class SyntheticInteraction:
    space: str
    seed: int
    spe: np.ndarray
    v_t1: np.ndarray
    v_t0: np.ndarray
    lines: list = field(default_factory=list)   # (a,b,c,d,J,V) with 1-based snt orbit indices


def make_interaction(space: str, seed: int, spe_range=(-4.0, 0.0), tbme_sigma=1.5,
                     base: "SyntheticInteraction | None" = None, perturb_sigma=0.0):
    """Draw a seeded isospin interaction and convert it to snt pp/nn/pn lines.

    With `base`, SPEs are the base's and every isospin TBME is the base value
    plus N(0, perturb_sigma) drawn from this seed (interaction variants).
    """
    orbits = SPACES[space]
    lab1, lab0 = isospin_labels(orbits)
    rng = np.random.Generator(np.random.PCG64(seed))
    if base is None:
        spe = rng.uniform(spe_range[0], spe_range[1], size=len(orbits))
        v1 = rng.normal(0.0, tbme_sigma, size=len(lab1))
        v0 = rng.normal(0.0, tbme_sigma, size=len(lab0))
    else:
        spe = base.spe.copy()
        v1 = base.v_t1 + rng.normal(0.0, perturb_sigma, size=len(lab1))
        v0 = base.v_t0 + rng.normal(0.0, perturb_sigma, size=len(lab0))
    lp = len(orbits)
    lines = []
    for (a, b, c, d, J), v in zip(lab1, v1):           # pp
        lines.append((a, b, c, d, J, float(v)))
    for (a, b, c, d, J), v in zip(lab1, v1):           # nn
        lines.append((a + lp, b + lp, c + lp, d + lp, J, float(v)))
    pn_pairs = []
    for a in range(1, lp + 1):
        for b in range(1, lp + 1):
            ja, jb = orbits[a - 1][2], orbits[b - 1][2]
            for J in range(abs(ja - jb) // 2, (ja + jb) // 2 + 1):
                pn_pairs.append((a, b, J))
    for x in range(len(pn_pairs)):
        for y in range(x, len(pn_pairs)):
            a, b, J = pn_pairs[x]
            c, d, J2 = pn_pairs[y]
            if J != J2:
                continue
            i1, f1 = _pn_factor(orbits, (a, b, c, d, J), lab1, 1)
            i0, f0 = _pn_factor(orbits, (a, b, c, d, J), lab0, 0)
            val = 0.0
            if i1 >= 0:
                val += v1[i1] * f1
            if i0 >= 0:
                val += v0[i0] * f0
            lines.append((a, b + lp, c, d + lp, J, float(val)))
    return SyntheticInteraction(space, seed, spe, v1, v0, lines)


def write_snt(path: str, inter: SyntheticInteraction, mass: int, comment: str = "") -> str:
    """Write `inter` as an snt file (values `%.10e`; massop 0 = no scaling)."""
    orbits = SPACES[inter.space]
    cp, cn = CORES[inter.space]
    lp = len(orbits)
    out = []
    out.append(f"! synthetic isospin interaction space={inter.space} seed={inter.seed} {comment}".rstrip())
    out.append(f"  {lp}  {lp}  {cp}  {cn}")
    for i, (n, l, j) in enumerate(orbits, start=1):
        out.append(f"  {i}  {n}  {l}  {j}  -1")
    for i, (n, l, j) in enumerate(orbits, start=1):
        out.append(f"  {i + lp}  {n}  {l}  {j}  1")
    out.append(f"  {2 * lp}  0")
    for i in range(lp):
        out.append(f"  {i + 1}  {i + 1}  {inter.spe[i]:.10e}")
    for i in range(lp):
        out.append(f"  {i + 1 + lp}  {i + 1 + lp}  {inter.spe[i]:.10e}")
    out.append(f"  {len(inter.lines)}  0  {mass}")
    for a, b, c, d, J, v in inter.lines:
        out.append(f"  {a}  {b}  {c}  {d}  {J}  {v:.10e}")
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    return path


BASE_SEED = 20261203


def config_snt(path: str, space: str, zv: int, nv: int, seed: int, spe_range, sigma) -> tuple:
    """Write the snt of one synthetic configuration; returns (path, nucleus)."""
    nuc = nucleus_name(space, zv, nv)
    mass = CORES[space][0] + CORES[space][1] + zv + nv
    inter = make_interaction(space, seed, spe_range, sigma)
    write_snt(path, inter, mass)
    return path, nuc


def variant_snts(directory: str, n_variants: int, space="pf", zv=4, nv=4, seed=BASE_SEED + 5,
                 spe_range=(-8.0, 0.0), sigma=1.0, perturb_sigma=0.05) -> tuple:
    """Interaction variants sharing one basis: a base isospin interaction (seed)
    plus, per variant v, N(0, perturb_sigma) on every TBME drawn from seed
    `seed * 1000 + v`; SPEs from the base.  Returns (paths, nucleus)."""
    nuc = nucleus_name(space, zv, nv)
    mass = CORES[space][0] + CORES[space][1] + zv + nv
    base = make_interaction(space, seed, spe_range, sigma)
    paths = []
    for v in range(n_variants):
        inter = make_interaction(space, seed * 1000 + v, base=base, perturb_sigma=perturb_sigma)
        paths.append(write_snt(os.path.join(directory, f"variant_{v:03d}.snt"), inter, mass,
                               comment=f"variant={v} base_seed={seed}"))
    return paths, nuc


@dataclass
class ServingRequest:
    index: int
    space: str
    zv: int
    nv: int
    num_ev: int
    arrival_s: float     # offset from the start of the stream
    snt: str
    nucleus: str
    seed: int


def serving_stream(directory: str, n_requests: int, seed=BASE_SEED * 10 + 4, rate_per_s=50.0,
                   spaces=("sd", "pf"), particles=(2, 6), num_ev=(1, 10), spe_range=(-6.0, 0.0), sigma=1.2,
                   write=True) -> list:
    """A seeded stream of eigensolve requests: space uniform over `spaces`,
    proton and neutron valence numbers uniform in `particles` (inclusive), k
    uniform in `num_ev`, Poisson arrivals (exponential gaps, mean 1/rate), and a
    per-request interaction drawn from seed `seed * 100000 + i`.  2M is left to
    the main_sm default (A % 2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    t = 0.0
    for i in range(n_requests):
        space = spaces[int(rng.integers(len(spaces)))]
        zv = int(rng.integers(particles[0], particles[1] + 1))
        nv = int(rng.integers(particles[0], particles[1] + 1))
        k = int(rng.integers(num_ev[0], num_ev[1] + 1))
        t += float(rng.exponential(1.0 / rate_per_s))
        rseed = seed * 100000 + i
        path = os.path.join(directory, f"request_{i:04d}.snt")
        nuc = nucleus_name(space, zv, nv)
        if write:
            config_snt(path, space, zv, nv, rseed, spe_range, sigma)
        out.append(ServingRequest(i, space, zv, nv, k, t, path, nuc, rseed))
    return out
