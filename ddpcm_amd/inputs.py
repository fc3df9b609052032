"""Input.txt reader/writer and the seeded synthetic-molecule generator.

The file format is the reference driver's (main.f90:124-155): eight scalar
lines (iprint, nproc, lmax, ngrid, iconv, igrad, eps, eta), the atom count n,
then n records ``q x y z r`` in angstrom.  Each read is list-directed, so only
the leading values of a line count and trailing comments are ignored; only the
first n records are read even when more follow (the root Input.txt of the
reference declares n=10 over 24 036 records).

Coordinates and radii are converted to bohr with the reference's constant
(main.f90:99,149-153).
"""
from __future__ import annotations

import dataclasses
import math
import re

import numpy as np

TOANG = 0.52917721092          # main.f90:99
TOKCAL = 627.509469            # main.f90:99

_SPLIT = re.compile(r"[\s,]+")


@dataclasses.dataclass
class DDInput:
    iprint: int
    nproc: int
    lmax: int
    ngrid: int
    iconv: int
    igrad: int
    eps: float
    eta: float
    charge: np.ndarray   # (n,)
    xyz: np.ndarray      # (n, 3) angstrom
    rvdw: np.ndarray     # (n,) angstrom

    @property
    def n(self) -> int:
        return int(self.charge.shape[0])

    def bohr(self):
        """(x, y, z, r) in bohr, exactly as main.f90:149-153 converts them."""
        tobohr = 1.0 / TOANG
        return (self.xyz[:, 0] * tobohr, self.xyz[:, 1] * tobohr,
                self.xyz[:, 2] * tobohr, self.rvdw * tobohr)


def _lead(line: str, k: int):
    toks = [t for t in _SPLIT.split(line.strip()) if t]
    if len(toks) < k:
        raise ValueError("Input.txt: expected %d values in line %r" % (k, line))
    return toks[:k]


def _num(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def read_input(path: str) -> DDInput:
    with open(path) as f:
        lines = f.read().splitlines()
    ints = [int(_lead(lines[i], 1)[0]) for i in (0, 1, 2, 3, 4, 5)]
    eps = _num(_lead(lines[6], 1)[0])
    eta = _num(_lead(lines[7], 1)[0])
    n = int(_lead(lines[8], 1)[0])
    rec = np.array([[_num(t) for t in _lead(lines[9 + i], 5)] for i in range(n)],
                   dtype=np.float64).reshape(n, 5)
    return DDInput(*ints, eps, eta, rec[:, 0].copy(), rec[:, 1:4].copy(),
                   rec[:, 4].copy())


def write_input(path: str, d: DDInput) -> None:
    with open(path, "w") as f:
        f.write("%d\n%d\n%d\n%d\n%d\n%d\n" % (d.iprint, d.nproc, d.lmax, d.ngrid,
                                             d.iconv, d.igrad))
        f.write("%.10g\n%.10g\n%d\n" % (d.eps, d.eta, d.n))
        for i in range(d.n):
            f.write("%.17g %.17g %.17g %.17g %.17g\n" % (
                d.charge[i], d.xyz[i, 0], d.xyz[i, 1], d.xyz[i, 2], d.rvdw[i]))


# --------------------------------------------------------------------------
# Synthetic molecules.
#
# Cubic lattice of 1.5 A spacing; ball radius R = (N / (0.1 * 4 pi / 3))^(1/3)
# A (about 0.1 atoms per A^3); a uniform random subset of the lattice sites in
# the ball, each coordinate jittered by U(-0.3, 0.3) A; elements H/C/N/O drawn
# with p = 0.5/0.3/0.1/0.1 and Bondi radii (1.20, 1.70, 1.55, 1.52 A) x 1.1;
# charges U[-0.6, 0.6] shifted to zero net charge.

BONDI = {"H": 1.20, "C": 1.70, "N": 1.55, "O": 1.52}
ELEMENTS = ("H", "C", "N", "O")
ELEMENT_P = (0.5, 0.3, 0.1, 0.1)


def synthetic_molecule(n: int, seed: int, spacing: float = 1.5,
                       density: float = 0.1, jitter: float = 0.3):
    """Returns (charge (n,), xyz (n,3) angstrom, rvdw (n,) angstrom)."""
    rng = np.random.default_rng(seed)
    radius = (n / (density * 4.0 * math.pi / 3.0)) ** (1.0 / 3.0)
    while True:
        k = int(math.ceil(radius / spacing)) + 1
        ax = np.arange(-k, k + 1) * spacing
        g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
        g = g[np.einsum("ij,ij->i", g, g) <= radius * radius]
        if g.shape[0] >= n:
            break
        radius *= 1.1
    pick = rng.choice(g.shape[0], size=n, replace=False)
    xyz = g[pick] + rng.uniform(-jitter, jitter, size=(n, 3))
    elem = rng.choice(len(ELEMENTS), size=n, p=ELEMENT_P)
    rvdw = np.array([BONDI[ELEMENTS[e]] for e in elem]) * 1.1
    q = rng.uniform(-0.6, 0.6, size=n)
    q -= q.mean()
    return q, xyz, rvdw


def synthetic_input(n: int, seed: int, lmax: int, ngrid: int, iconv: int = 8,
                    igrad: int = 1, eps: float = 78.39, eta: float = 0.1,
                    iprint: int = 0, nproc: int = 1) -> DDInput:
    q, xyz, r = synthetic_molecule(n, seed)
    return DDInput(iprint, nproc, lmax, ngrid, iconv, igrad, eps, eta, q, xyz, r)


# The benchmark configurations (atom count, lmax, grid, forces) with fixed seeds.
CONFIGS = {
    "c1": dict(n=12, seed=12, lmax=6, ngrid=302, igrad=0),
    "c2": dict(n=2000, seed=2000, lmax=8, ngrid=590, igrad=1),
    "c3": dict(n=20000, seed=20000, lmax=10, ngrid=1202, igrad=1),
}


def config_input(name: str, **over) -> DDInput:
    c = dict(CONFIGS[name])
    c.update(over)
    return synthetic_input(**c)


def batch_sizes(count: int = 512, lo: int = 20, hi: int = 800, seed: int = 4000):
    """Atom counts drawn log-uniform in [lo, hi] for the batched workload."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), size=count))).astype(int)


def batch_inputs(count: int = 512, lo: int = 20, hi: int = 800, seed: int = 4000,
                 lmax: int = 6, ngrid: int = 302, igrad: int = 1) -> list:
    """The batched workload: `count` molecules of batch_sizes() atoms, molecule
    q drawn with seed + 1 + q."""
    return [synthetic_input(int(n), seed + 1 + q, lmax, ngrid, igrad=igrad)
            for q, n in enumerate(batch_sizes(count, lo, hi, seed))]


def trajectory(n: int = 1000, frames: int = 2000, seed: int = 1000,
               sigma: float = 0.05):
    """Yields (charge, xyz, rvdw) per frame: a synthetic molecule whose atoms
    take seeded Gaussian steps of sigma angstrom per frame."""
    q, xyz, r = synthetic_molecule(n, seed)
    rng = np.random.default_rng(seed + 1)
    for _ in range(frames):
        yield q, xyz.copy(), r
        xyz = xyz + rng.normal(0.0, sigma, size=xyz.shape)
