"""RearrangementInt (CalcFunctions.cpp:469-545) on the device pair sums.

The correction for local rearrangements of a block's cell atoms: pairs of
cell atoms closer than each atom's furthest listed neighbour (plus Rcorr)
are taken out of the intensity and the listed neighbour distances are put
in, both weighted by the number of cell replicas (copies x cells x symmetry
images), then I = (I0 Ntot + I_rearr) / sqrt(N) with N the pair count after
the exchange.  The pair sums run on the GPU (``Context.cutoff_pairsum`` and
``Context.pairlist_pairsum``); the weighting, the pair bookkeeping, the
partial-intensity layout and the polarisation factor are host FP64 in the
reference's order.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional, Sequence

import numpy as np

PI = 3.1415926535897932384626433832795028841971693993  # typedefs.h:58


@dataclass
class RearrBlock:
    """A block's cell atoms (element-major, as after block::sortAtoms) with
    neighbour lists.  neighbours[i] is a (k, 4) array of (element, x, y, z)
    rows: the list read for cell atom i (block::readRearrangement, which is
    indexed by the atom's position in the input file while r_atom is in
    sorted order -- the reference pairs them by index as well).  The lists
    may also be given flat: neighbours a (K, 4) array of all rows in atom
    order and nb_count[i] the number of rows of atom i.
    coeff = Ncopy * Ncell.x * Ncell.y * Ncell.z * Nsymm."""
    r_atom: np.ndarray
    el: np.ndarray
    neighbours: Sequence[np.ndarray]
    rcorr: float
    coeff: int
    rearrangement: bool = True
    nb_count: Optional[np.ndarray] = None


def _flat(neighbours, nb_count=None):
    """All neighbour rows (k, 4) in atom order and the atom of each."""
    if nb_count is not None:
        counts = np.asarray(nb_count, np.int64)
        return (np.asarray(neighbours, np.float64).reshape(-1, 4),
                np.repeat(np.arange(counts.size), counts))
    counts = np.array([len(nb) for nb in neighbours], dtype=np.int64)
    rows = [np.asarray(nb, np.float64).reshape(-1, 4) for nb in neighbours if len(nb)]
    flat = np.concatenate(rows) if rows else np.zeros((0, 4))
    return flat, np.repeat(np.arange(len(neighbours)), counts)


def _sqr(r_atom, flat, idx):
    # (rAtom[i] - rAtomNeighb[i][k]).sqr(): (dx^2 + dy^2) + dz^2
    d = r_atom[idx] - flat[:, 1:4]
    return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]


def _prepare(r_atom, el, neighbours, rcorr, n_el, nb_count=None):
    """(rmax2, neighbour distances, neighbour types) of one block."""
    r_atom = np.asarray(r_atom, np.float64)
    flat, idx = _flat(neighbours, nb_count)
    m = np.zeros(len(r_atom))
    if not idx.size:
        return m + rcorr * rcorr, np.zeros(0), np.zeros(0, np.uint32)
    s = _sqr(r_atom, flat, idx)
    np.maximum.at(m, idx, s)
    t = np.asarray(el, np.int64)[idx] * n_el + flat[:, 0].astype(np.int64)
    return m + rcorr * rcorr, np.sqrt(s), t.astype(np.uint32)


def rmax2(r_atom, neighbours, rcorr, nb_count=None) -> np.ndarray:
    """Rmax2(rAtom[i], rAtomNeighb[i]) + SQR(Rcorr) (CalcFunctions.cpp:244-250, 490):
    the largest squared neighbour distance (0 without neighbours)."""
    return _prepare(r_atom, np.zeros(len(r_atom), np.int64), neighbours, rcorr, 1, nb_count)[0]


def neighbour_list(r_atom, el, neighbours, n_el, nb_count=None):
    """Distances |r_i - r_n| (mag(), CalcFunctions.cpp:510) and types
    el_i * n_el + el_n of all listed neighbours, atom by atom."""
    return _prepare(r_atom, el, neighbours, 0.0, n_el, nb_count)[1:]


def polar_1d(I, q, lam):
    """PolarFactor1D (CalcFunctions.cpp:460-468), in place."""
    s = q * (lam * 0.25 / PI)
    c = 1. - 2. * (s * s)
    I *= 0.5 * (1. + c * c)


def rearrangement_intensity(I0, q, blocks: Sequence[RearrBlock], n_el: int, n_tot: int,
                            source: str, ff=None, sl=None, partial: bool = False,
                            polar: bool = False, lam: float = 0.0,
                            cutoff_sums: Optional[Callable] = None,
                            list_sums: Optional[Callable] = None, ctx=None) -> np.ndarray:
    """RearrangementInt: returns the updated I0 (a new array).

    I0 has nq values, or (1 + nblocks (nblocks + 1) / 2) nq with partial
    intensities (total first, then block pairs).  source "xray" weights
    with ff[el, q], "neutron" with sl[el].  The pair sums come from ``ctx``
    (a Context) unless ``cutoff_sums(xyz, el, n_el, rmax2, q)`` and
    ``list_sums(r, types, ntypes, q)`` are given."""
    if cutoff_sums is None:
        cutoff_sums = ctx.cutoff_pairsum
    if list_sums is None:
        list_sums = ctx.pairlist_pairsum
    q = np.ascontiguousarray(q, np.float64)
    nq = q.size
    I0 = np.array(I0, dtype=np.float64)
    nb_ = len(blocks)
    xray = source == "xray"
    # long long N = SQR(Ntot): the product is formed in unsigned int
    N = (int(n_tot) * int(n_tot)) & 0xFFFFFFFF
    I = np.zeros((nb_ + 1) * nq if partial else nq)
    for B in blocks:
        iqst = nq if partial else 0  # the same slot for every block, CalcFunctions.cpp:478-479
        if not B.rearrangement:
            continue
        coeff = int(B.coeff) & 0xFFFFFFFF
        r = np.ascontiguousarray(B.r_atom, np.float64)
        el = np.asarray(B.el, np.uint32)
        r2, rr, tt = _prepare(r, el, B.neighbours, B.rcorr, n_el, B.nb_count)
        S_cut, npairs = cutoff_sums(np.ascontiguousarray(r.T), el, n_el, r2, q)
        S_nb = list_sums(rr, tt, n_el * n_el, q)
        nnb = np.bincount(tt // n_el, minlength=n_el) if tt.size else np.zeros(n_el, np.int64)
        nb_types = np.bincount(tt, minlength=n_el * n_el) if tt.size else \
            np.zeros(n_el * n_el, np.int64)
        present = [int(e) for e in np.unique(el)]  # block::id: ascending ensemble elements
        seg = slice(iqst, iqst + nq)
        for a in present:
            for b in present:
                t = a * n_el + b
                N -= coeff * int(npairs[t])
                w = (coeff * ff[a] * ff[b]) if xray else (coeff * sl[a] * sl[b])
                I[seg] -= w * S_cut[t]
            for b in range(n_el):
                t = a * n_el + b
                if not nb_types[t]:
                    continue
                w = (coeff * ff[a] * ff[b]) if xray else (coeff * sl[a] * sl[b])
                I[seg] += w * S_nb[t]
            N += coeff * int(nnb[a])
    if partial:
        iqst = nq
        for iB in range(nb_):
            i0 = nq * (1 + iB * (2 * nb_ - iB + 1) // 2)
            if polar:
                polar_1d(I[iqst:iqst + nq], q, lam)
            I[:nq] += I[iqst:iqst + nq]
            I0[i0:i0 + nq] = (I0[i0:i0 + nq] * n_tot + I[iqst:iqst + nq]) / np.sqrt(float(N))
            iqst += nq
    elif polar:
        polar_1d(I, q, lam)
    I0[:nq] = (I0[:nq] * n_tot + I[:nq]) / np.sqrt(float(N))
    return I0
