"""Seeded synthetic atomic ensembles and form factors.

These stand in for the reference's ensemble builder (CalcFunctions.cpp:251-450,
block::calcAtoms), whose random generator is clock-seeded
(CalcFunctions.cpp:259) and so cannot reproduce a run.  Every builder here
draws from numpy's PCG64 with an explicit seed, stores positions as float32
(the precision the GPU kernels consume) and returns atoms grouped
element-major, the order block::sortAtoms (CalcFunctions.cpp:159-196)
produces.

A structure is returned as ``Ensemble(xyz, el_count, names)`` with ``xyz`` a
(3, N) float32 array (rows x, y, z).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

FOUR_PI = 4.0 * np.pi


@dataclass
class Ensemble:
    xyz: np.ndarray                      # (3, N) float32, element-major
    el_count: List[int]                  # atoms per element
    names: List[str] = field(default_factory=list)

    @property
    def n(self) -> int:
        return int(self.xyz.shape[1])

    @property
    def n_el(self) -> int:
        return len(self.el_count)

    def xyz64(self) -> np.ndarray:
        """Positions promoted exactly to float64, the oracle's input."""
        return np.ascontiguousarray(self.xyz, dtype=np.float64)

    def element(self, e: int) -> np.ndarray:
        off = int(sum(self.el_count[:e]))
        return self.xyz[:, off:off + self.el_count[e]]


def _concat(parts: Sequence[np.ndarray], names: Sequence[str]) -> Ensemble:
    counts = [int(p.shape[0]) for p in parts]
    xyz = np.concatenate([p.reshape(-1, 3) for p in parts], axis=0) if parts else np.zeros((0, 3))
    return Ensemble(np.ascontiguousarray(xyz.T.astype(np.float32)), counts, list(names))


def q_mesh(qmin: float, qmax: float, n: int) -> np.ndarray:
    """q_k = qmin + k*(qmax-qmin)/n, k < n: the endpoint-exclusive mesh of
    ReadXML.cpp:220-222."""
    dq = (qmax - qmin) / n
    return np.array([qmin + k * dq for k in range(n)], dtype=np.float64)


def fcc_sites(a: float, nmax: int) -> np.ndarray:
    """fcc lattice sites with cell indices in [-nmax, nmax]^3."""
    r = np.arange(-nmax, nmax + 1)
    cx, cy, cz = np.meshgrid(r, r, r, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1).astype(np.float64)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    return ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a


def bcc_sites(a: float, nmax: int) -> np.ndarray:
    r = np.arange(-nmax, nmax + 1)
    cx, cy, cz = np.meshgrid(r, r, r, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1).astype(np.float64)
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0.5]])
    return ((cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)) * a


def fcc_sphere(a: float = 4.078, radius: float = 20.0, sigma: float = 0.05,
               seed: int = 1, name: str = "Au") -> Ensemble:
    """Single-element fcc lattice cut to a sphere centred on a lattice site,
    Gaussian displacement sigma per coordinate (workload: 2 nm-radius fcc
    sphere, about 2,000 atoms)."""
    nmax = int(np.ceil(radius / a)) + 1
    s = fcc_sites(a, nmax)
    s = s[np.einsum("ij,ij->i", s, s) <= radius * radius]
    rng = np.random.Generator(np.random.PCG64(seed))
    s = s + rng.normal(0.0, sigma, size=s.shape)
    return _concat([s], [name])


def rocksalt_cube(a: float = 5.64, ncell: int = 50, vacancy: float = 0.02,
                  sigma: float = 0.08, seed_vac: int = 3, seed_jit: int = 4,
                  names=("Na", "Cl")) -> Ensemble:
    """Two-element rock-salt cube of ncell^3 conventional cells (8*ncell^3
    sites), Bernoulli vacancies and Gaussian jitter (workload: 1,000,000-site
    rock-salt cube, a=5.64 A, 2% vacancies, sigma=0.08 A)."""
    r = np.arange(ncell, dtype=np.float64)
    cx, cy, cz = np.meshgrid(r, r, r, indexing="ij")
    cells = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1)
    fcc = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    shift = np.array([0.5, 0.0, 0.0])
    na = ((cells[:, None, :] + fcc[None]).reshape(-1, 3)) * a
    cl = ((cells[:, None, :] + (fcc + shift)[None]).reshape(-1, 3)) * a
    rv = np.random.Generator(np.random.PCG64(seed_vac))
    keep_na = rv.random(na.shape[0]) >= vacancy
    keep_cl = rv.random(cl.shape[0]) >= vacancy
    na, cl = na[keep_na], cl[keep_cl]
    rj = np.random.Generator(np.random.PCG64(seed_jit))
    na = na + rj.normal(0.0, sigma, size=na.shape)
    cl = cl + rj.normal(0.0, sigma, size=cl.shape)
    return _concat([na, cl], list(names))


def disordered_cubic(a: float = 2.5, n: int = 37, jitter: float = 0.6,
                     fractions=(0.6, 0.3, 0.1), natoms: int | None = 50000,
                     seed_jit: int = 6, seed_species: int = 7,
                     names=("Fe", "Ni", "Cr")) -> Ensemble:
    """Simple-cubic lattice with uniform jitter in [-j, j] per coordinate and
    species drawn with the given fractions, centred on the origin (workload:
    50,000-atom disordered solid, 3 elements 60/30/10%)."""
    r = np.arange(n, dtype=np.float64)
    cx, cy, cz = np.meshgrid(r, r, r, indexing="ij")
    s = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1) * a
    if natoms is not None:
        s = s[:natoms]
    rj = np.random.Generator(np.random.PCG64(seed_jit))
    s = s + rj.uniform(-jitter, jitter, size=s.shape)
    s = s - s.mean(axis=0)
    rs = np.random.Generator(np.random.PCG64(seed_species))
    u = rs.random(s.shape[0])
    edges = np.cumsum(fractions)
    species = np.searchsorted(edges, u, side="right")
    species = np.minimum(species, len(fractions) - 1)
    parts = [s[species == e] for e in range(len(fractions))]
    return _concat(parts, list(names[:len(fractions)]))


def gaussian_ff_coeffs(n_el: int, seed: int = 2) -> np.ndarray:
    """(n_el, 9) rows a1..a4, b1..b4, c with a_k~U[0.5,20], b_k~U[0.1,100] A^2,
    c~U[0,2] (a 4-Gaussian x-ray form factor)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((n_el, 9))
    for e in range(n_el):
        out[e, 0:4] = rng.uniform(0.5, 20.0, 4)
        out[e, 4:8] = rng.uniform(0.1, 100.0, 4)
        out[e, 8] = rng.uniform(0.0, 2.0)
    return out


def gaussian_ff(q: np.ndarray, coeffs: np.ndarray) -> np.ndarray:
    """f_el(q) = sum_k a_k exp(-b_k (q/4pi)^2) + c, evaluated in float64.
    Returns (n_el, len(q))."""
    q = np.asarray(q, dtype=np.float64)
    s2 = (q / FOUR_PI) ** 2
    coeffs = np.atleast_2d(coeffs)
    out = np.empty((coeffs.shape[0], q.size))
    for e, c in enumerate(coeffs):
        out[e] = sum(c[k] * np.exp(-c[4 + k] * s2) for k in range(4)) + c[8]
    return out


def random_cluster(n_per_el: Sequence[int], box: float, min_dist: float,
                   seed: int, names: Sequence[str] | None = None) -> Ensemble:
    """Random atoms in a cube with a minimum separation (rejection sampling),
    for small irregular test cases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    total = int(sum(n_per_el))
    pts: list = []
    while len(pts) < total:
        p = rng.uniform(-box / 2, box / 2, 3)
        if all(np.sum((p - o) ** 2) >= min_dist ** 2 for o in pts):
            pts.append(p)
    pts = np.array(pts)
    parts, o = [], 0
    for c in n_per_el:
        parts.append(pts[o:o + c])
        o += c
    names = list(names) if names else [f"E{e}" for e in range(len(n_per_el))]
    return _concat(parts, names)


def particle(n_target: int, n_el: int, shape: str, a: float, sigma: float,
             rng: np.random.Generator) -> Ensemble:
    """One nanoparticle of about n_target atoms.

    shape: "fcc" / "bcc" (sphere cut from the lattice), "cube" (fcc cube) or
    "core-shell" (fcc sphere, element 0 inside 60% of the radius, the other
    elements drawn at random outside it).  Species of the non-core shapes are
    drawn with Dirichlet(1) fractions.  Gaussian jitter sigma per coordinate."""
    if shape == "bcc":
        r = (3.0 * n_target * a ** 3 / (8.0 * np.pi)) ** (1.0 / 3.0)
        s = bcc_sites(a, int(np.ceil(r / a)) + 1)
        s = s[np.einsum("ij,ij->i", s, s) <= r * r]
    elif shape == "cube":
        half = 0.5 * (n_target * a ** 3 / 4.0) ** (1.0 / 3.0)
        s = fcc_sites(a, int(np.ceil(half / a)) + 1)
        s = s[np.all(np.abs(s) <= half, axis=1)]
    else:
        r = (3.0 * n_target * a ** 3 / (16.0 * np.pi)) ** (1.0 / 3.0)
        s = fcc_sites(a, int(np.ceil(r / a)) + 1)
        s = s[np.einsum("ij,ij->i", s, s) <= r * r]
    s = s + rng.normal(0.0, sigma, size=s.shape)
    if shape == "core-shell" and n_el > 1:
        rr = np.sqrt(np.einsum("ij,ij->i", s, s))
        core = rr <= 0.6 * rr.max()
        species = np.where(core, 0, 1 + rng.integers(0, n_el - 1, size=s.shape[0]))
    else:
        frac = rng.dirichlet(np.ones(n_el))
        species = np.minimum(np.searchsorted(np.cumsum(frac), rng.random(s.shape[0]), side="right"),
                             n_el - 1)
    parts = [s[species == e] for e in range(n_el)]
    if any(p.shape[0] == 0 for p in parts):     # keep every element present
        order = rng.permutation(s.shape[0])
        species = np.arange(s.shape[0])[order] % n_el
        parts = [s[species == e] for e in range(n_el)]
    return _concat(parts, [f"E{e}" for e in range(n_el)])


def particle_ensemble(count: int = 1000, seed: int = 9, nmin: int = 500,
                      nmax: int = 200000) -> List[Ensemble]:
    """An ensemble of nanoparticles: N log-uniform in [nmin, nmax], 1-4
    elements, shapes fcc/bcc sphere, fcc cube and core-shell, sigma = 0.05 A
    (workload: 1,000 particles of 500-200,000 atoms)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shapes = ("fcc", "bcc", "cube", "core-shell")
    out = []
    for _ in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        n_el = int(rng.integers(1, 5))
        shape = shapes[int(rng.integers(0, 4))]
        a = float(rng.uniform(2.8, 3.4) if shape == "bcc" else rng.uniform(3.5, 4.2))
        out.append(particle(n, n_el, shape, a, 0.05, rng))
    return out


def c4_particles(count: int = 256, seed: int = 11, nmin: int = 500,
                 nmax: int = 8000) -> List[Ensemble]:
    """Two-element fcc and core-shell spheres, N log-uniform in [nmin, nmax],
    a in [3.6, 4.1] A, so every interatomic distance stays below 65.5 A
    (workload: 256 particles x 3 pair types on one 16,384-bin grid at
    0.004 A)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        shape = ("fcc", "core-shell")[int(rng.integers(0, 2))]
        out.append(particle(n, 2, shape, float(rng.uniform(3.6, 4.1)), 0.05, rng))
    return out


def uniform_rotations(n: int, seed: int = 8) -> np.ndarray:
    """n rotation matrices (n, 3, 3) uniform on SO(3), from uniform unit
    quaternions (Shoemake's subgroup algorithm)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    w, x = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2)
    y, z = b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def q_plane(qmax: float = 8.0, n: int = 512) -> np.ndarray:
    """(n*n, 3) Cartesian q vectors, qx, qy on n points from -qmax to qmax
    (exactly symmetric about 0), qz = 0."""
    g = (np.arange(n) - (n - 1) / 2.0) * (2.0 * qmax / (n - 1))
    QX, QY = np.meshgrid(g, g, indexing="ij")
    return np.stack([QX.ravel(), QY.ravel(), np.zeros(QX.size)], axis=1)
