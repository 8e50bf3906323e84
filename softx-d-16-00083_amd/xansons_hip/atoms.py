"""Ensemble builder on the device: structural blocks -> atoms (xs_atoms_*).

A :class:`Block` holds what block::calcAtoms (CalcFunctions.cpp:251-397)
reads from one structural block after ReadConfig: absolute cell-atom
positions sorted element-major, symmetry operations, cell vectors and cell
counts, copies, cut-offs and displacement parameters.  :class:`Atoms` builds
the ensemble of a list of blocks on the GPU; its coordinates are
element-major SoA, the layout every calculator of this package takes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import XS_DEVICE_PTRS, Context, _check, lib


def _num(v) -> float:
    return float.fromhex(v) if isinstance(v, str) else float(v)


def _arr(v, shape) -> np.ndarray:
    a = np.array([_num(x) for x in np.ravel(v)], dtype=np.float64) if len(np.ravel(v)) else \
        np.zeros(0)
    return a.reshape(shape)


@dataclass
class Block:
    """One structural block (see the module docstring).

    r_atom (natom, 3) and el (natom,) are required; the rest default to one
    cell, the identity symmetry operation, one unrotated copy at the origin,
    occupancy 1, no displacements and no cut-offs."""
    r_atom: np.ndarray
    el: np.ndarray
    occ: Optional[np.ndarray] = None
    dev: Optional[np.ndarray] = None
    p_symm: Optional[np.ndarray] = None  # (nsymm, 3, 3): rows R1, R2, R3
    r_symm: Optional[np.ndarray] = None  # (nsymm, 3) absolute translations
    cell: Optional[np.ndarray] = None    # (3, 3): rows a, b, c
    ncell: Sequence[int] = (1, 1, 1)
    r_copy: Optional[np.ndarray] = None  # (ncopy, 3)
    cf: Optional[np.ndarray] = None      # (ncopy, 3, 3): rotation rows
    r_mean: Optional[np.ndarray] = None
    rcut: float = 0.0
    rcut_copies: float = 0.0
    dev_mol: float = 0.0
    mol_rotation: bool = False

    def __post_init__(self):
        self.r_atom = np.ascontiguousarray(self.r_atom, dtype=np.float64).reshape(-1, 3)
        n = self.r_atom.shape[0]
        self.el = np.ascontiguousarray(self.el, dtype=np.uint32).reshape(n)
        self.occ = np.ones(n) if self.occ is None else np.ascontiguousarray(self.occ, np.float64)
        self.dev = np.zeros(n) if self.dev is None else np.ascontiguousarray(self.dev, np.float64)
        self.p_symm = (np.eye(3)[None] if self.p_symm is None else
                       np.ascontiguousarray(self.p_symm, np.float64).reshape(-1, 3, 3))
        self.r_symm = (np.zeros((self.p_symm.shape[0], 3)) if self.r_symm is None else
                       np.ascontiguousarray(self.r_symm, np.float64).reshape(-1, 3))
        self.cell = np.eye(3) if self.cell is None else \
            np.ascontiguousarray(self.cell, np.float64).reshape(3, 3)
        self.ncell = tuple(int(v) for v in self.ncell)
        self.r_copy = np.zeros((1, 3)) if self.r_copy is None else \
            np.ascontiguousarray(self.r_copy, np.float64).reshape(-1, 3)
        self.cf = (np.eye(3)[None].repeat(self.r_copy.shape[0], 0) if self.cf is None else
                   np.ascontiguousarray(self.cf, np.float64).reshape(-1, 3, 3))
        self.r_mean = np.zeros(3) if self.r_mean is None else \
            np.ascontiguousarray(self.r_mean, np.float64).reshape(3)
        if self.r_symm.shape[0] != self.p_symm.shape[0] or self.cf.shape[0] != self.r_copy.shape[0]:
            raise ValueError("Block: symmetry or copy arrays disagree in length")

    @property
    def natom(self) -> int:
        return self.r_atom.shape[0]

    @property
    def nsymm(self) -> int:
        return self.p_symm.shape[0]

    @property
    def ncopy(self) -> int:
        return self.r_copy.shape[0]

    def candidates(self) -> int:
        """Upper bound on the atoms this block places."""
        return self.natom * self.nsymm * self.ncopy * int(np.prod(self.ncell))

    @classmethod
    def from_dict(cls, d: dict) -> "Block":
        """From a dict of the xs_block_desc fields; numbers may be floats or
        C99 hex-float strings."""
        natom, nsymm, ncopy = int(d["natom"]), int(d["nsymm"]), int(d["ncopy"])
        return cls(r_atom=_arr(d["r_atom"], (natom, 3)), el=np.asarray(d["el"], np.uint32),
                   occ=_arr(d["occ"], (natom,)), dev=_arr(d["dev"], (natom,)),
                   p_symm=_arr(d["p_symm"], (nsymm, 3, 3)), r_symm=_arr(d["r_symm"], (nsymm, 3)),
                   cell=_arr(d["cell"], (3, 3)), ncell=tuple(d["ncell"]),
                   r_copy=_arr(d["r_copy"], (ncopy, 3)), cf=_arr(d["cf"], (ncopy, 3, 3)),
                   r_mean=_arr(d["r_mean"], (3,)), rcut=_num(d["rcut"]),
                   rcut_copies=_num(d["rcut_copies"]), dev_mol=_num(d["dev_mol"]),
                   mol_rotation=bool(d["mol_rotation"]))


class _BlockDesc(C.Structure):
    _fields_ = [("natom", C.c_uint), ("nsymm", C.c_uint), ("ncopy", C.c_uint),
                ("ncell", C.c_uint * 3), ("el", C.c_void_p), ("r_atom", C.c_void_p),
                ("occ", C.c_void_p), ("dev", C.c_void_p), ("p_symm", C.c_void_p),
                ("r_symm", C.c_void_p), ("cell", C.c_void_p), ("r_copy", C.c_void_p),
                ("cf", C.c_void_p), ("r_mean", C.c_double * 3), ("rcut", C.c_double),
                ("rcut_copies", C.c_double), ("dev_mol", C.c_double), ("mol_rotation", C.c_int)]


def _descs(blocks: Sequence[Block]):
    ds = (_BlockDesc * len(blocks))()
    for b, d in zip(blocks, ds):
        d.natom, d.nsymm, d.ncopy = b.natom, b.nsymm, b.ncopy
        d.ncell = (C.c_uint * 3)(*b.ncell)
        for k in ("el", "r_atom", "occ", "dev", "p_symm", "r_symm", "cell", "r_copy", "cf"):
            setattr(d, k, getattr(b, k).ctypes.data)
        d.r_mean = (C.c_double * 3)(*b.r_mean)
        d.rcut, d.rcut_copies, d.dev_mol = float(b.rcut), float(b.rcut_copies), float(b.dev_mol)
        d.mol_rotation = int(bool(b.mol_rotation))
    return ds


class Atoms:
    """The atoms of an ensemble of blocks, built on the device
    (xs_atoms_build).  ``seed`` fixes every random draw."""

    def __init__(self, ctx: Context, blocks: Sequence[Block], n_el: int, seed: int = 0):
        self.ctx = ctx
        self.h = C.c_void_p()
        ds = _descs(blocks)
        _check(lib().xs_atoms_build(ctx.h, C.addressof(ds), len(blocks), int(n_el),
                                    int(seed) & (2 ** 64 - 1), C.byref(self.h)))
        ctx._atoms.add(self)
        self.n_el, self.nblocks = int(n_el), len(blocks)
        self.el_count = np.zeros(self.n_el, dtype=np.uint32)
        self.block_el_count = np.zeros((self.nblocks, self.n_el), dtype=np.uint32)
        n = C.c_uint64(0)
        _check(lib().xs_atoms_info(self.h, self.el_count.ctypes.data,
                                   self.block_el_count.ctypes.data, C.byref(n)))
        self.n = int(n.value)

    def xyz(self) -> np.ndarray:
        """Element-major (3, n) float64 coordinates (host copy)."""
        out = np.zeros(3 * self.n)
        if self.n:
            _check(lib().xs_atoms_copy(self.h, out.ctypes.data, 0))
        return out.reshape(3, self.n)

    def copy_dev(self, ptr: int) -> None:
        """Enqueue a device-to-device copy of the SoA coordinates to ``ptr``."""
        _check(lib().xs_atoms_copy(self.h, C.c_void_p(ptr), XS_DEVICE_PTRS))

    def close(self):
        if self.h:
            lib().xs_atoms_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
