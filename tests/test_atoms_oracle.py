"""The oracle's ensemble builder (or_calc_atoms, a restatement of
block::calcAtoms, CalcFunctions.cpp:251-397) against the reference's own
builder: exact atoms on the deterministic golden cases (tests/golden/
make_atoms_golden.py), distributions on the random ones, and Philox4x64-10
against numpy's independent implementation."""
import json
import os
import sys

import numpy as np
import pytest

from conftest import REPO

sys.path.insert(0, os.path.join(REPO, "oracle"))
import xs_oracle as O  # noqa: E402
from xansons_hip.atoms import Block  # noqa: E402

GOLDEN = os.path.join(REPO, "tests", "golden")
DETERMINISTIC = ["rocksalt", "hex_copies", "shared_sites", "single_cell"]


def load(case):
    g = np.load(os.path.join(GOLDEN, f"atoms_{case}.npz"))
    blocks = [Block.from_dict(d) for d in json.loads(str(g["blocks_json"]))]
    return g, blocks


def element_major(el, blk, xyz, n_el):
    """Stable order by element: the reference's ra[el] lists."""
    order = np.argsort(el, kind="stable")
    counts = np.bincount(el, minlength=n_el)
    return xyz[order], counts


def xyz_lines(text):
    """Coordinate part of printAtoms' lines (IO.cpp:273-295), in file order."""
    lines = text.splitlines()
    assert int(lines[0]) == len(lines) - 2
    return [ln.split(" ", 1)[1] for ln in lines[2:]]


def fmt(xyz):
    # ostream << vect3d with ios::scientific, precision 6 (typedefs.h:146-153)
    return [f"{x:.6e}  {y:.6e}  {z:.6e}" for x, y, z in xyz]


def test_philox_matches_numpy():
    rng = np.random.default_rng(3)
    for _ in range(20):
        ctr = rng.integers(1, 2 ** 63, size=4, dtype=np.uint64)
        key = rng.integers(0, 2 ** 63, size=2, dtype=np.uint64)
        # numpy's Philox increments its counter before the first block
        prev = ctr.copy()
        prev[0] -= np.uint64(1)
        ref = np.random.Philox(counter=prev, key=key).random_raw(4)
        assert (O.philox4x64(ctr, key) == ref).all()


def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x64-10
    z = O.philox4x64([0] * 4, [0] * 2)
    assert [int(v) for v in z] == [0x16554d9eca36314c, 0xdb20fe9d672d0fdc,
                                   0xd7e772cee186176b, 0x7e68b68aec7ba23b]
    m = 2 ** 64 - 1
    f = O.philox4x64([m] * 4, [m] * 2)
    assert [int(v) for v in f] == [0x87b092c3013fe90b, 0x438c3c67be8d0224,
                                   0x9cc7d7c69cd777b6, 0xa09caebf594f0ba0]


@pytest.mark.parametrize("case", DETERMINISTIC)
def test_oracle_equals_reference_atoms(case):
    g, blocks = load(case)
    n_el = int(g["nel"])
    el, blk, xyz = O.calc_atoms(blocks, seed=12345)
    got, counts = element_major(el, blk, xyz, n_el)
    assert (counts == g["el_count"]).all()
    per_block = np.zeros((len(blocks), n_el), np.int64)
    np.add.at(per_block, (blk, el), 1)
    assert (per_block == g["block_el_count"]).all()
    # bit for bit
    assert got.shape == g["atoms"].shape
    assert (got.view(np.uint64) == g["atoms"].view(np.uint64)).all()
    # and the reference's printed file, character for character
    assert fmt(got) == xyz_lines(str(g["xyz_text"]))


@pytest.mark.parametrize("case", DETERMINISTIC)
def test_oracle_does_not_depend_on_seed(case):
    _, blocks = load(case)
    a = O.calc_atoms(blocks, seed=1)
    b = O.calc_atoms(blocks, seed=2)
    assert all((x == y).all() for x, y in zip(a, b))


def test_single_cell_merge_brute_force():
    """One cell, one symmetry operation, every occupancy 1: an atom is placed
    unless an earlier placed atom lies within 0.5 A of it
    (CalcFunctions.cpp:300-301, 340-346).  Restated with a dense numpy
    distance matrix, independent of the oracle's loops."""
    g, blocks = load("single_cell")
    B = blocks[0]
    r = B.r_atom
    d = r[:, None, :] - r[None, :, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    close = d2 < 0.25
    covered = np.zeros(B.natom, bool)
    placed = []
    for i in range(B.natom):
        if covered[i]:
            continue
        placed.append(i)
        covered |= close[i] & (np.arange(B.natom) >= i)
    assert len(placed) < B.natom  # the planted duplicates merge
    el = B.el[placed]
    order = np.argsort(el, kind="stable")
    want = r[np.asarray(placed)[order]]
    assert (want.view(np.uint64) == g["atoms"].view(np.uint64)).all()


def half_normal_stats(disp):
    r = np.linalg.norm(disp, axis=1)
    return r.mean(), (r ** 2).mean(), ((disp[:, 2] / r) ** 2).mean()


def test_random_occupancy_matches_reference_distribution():
    g, blocks = load("random_occ")
    ref = json.loads(str(g["stats_json"]))
    ncell = int(g["ncells"])
    sd = 0.1
    el, blk, xyz = O.calc_atoms(blocks, seed=777)
    counts = np.bincount(el, minlength=2)
    for c_ours, c_ref, p in zip(counts, g["el_count"], (0.35, 0.40)):
        s = np.sqrt(ncell * p * (1 - p))
        assert abs(c_ours - ncell * p) < 5 * s
        assert abs(int(c_ref) - ncell * p) < 5 * s
    disp = xyz - np.round(xyz / 4.0) * 4.0
    mr, mr2, mc2 = half_normal_stats(disp)
    n = len(disp)
    # |N(0, sd)| along (phi ~ U[0, 2pi), theta ~ U[0, pi)): E r = sd sqrt(2/pi),
    # E r^2 = sd^2, E cos^2(theta) = 1/2 (a uniform direction would give 1/3)
    exp = (sd * np.sqrt(2 / np.pi), sd ** 2, 0.5)
    se = (sd * np.sqrt(1 - 2 / np.pi), sd ** 2 * np.sqrt(2), np.sqrt(1 / 8))
    for ours, theirs, e, s in zip((mr, mr2, mc2), (ref["mean_r"], ref["mean_r2"], ref["mean_cos2"]),
                                  exp, se):
        assert abs(ours - e) < 5 * s / np.sqrt(n)
        assert abs(theirs - e) < 5 * s / np.sqrt(ref["n"])


def rotation_moment(bond, n=96):
    """E[R b b^T R^T] for R = Rz(a) Rx(b) Rz(g), a, b ~ U[0, 2pi), g ~ U[0, pi):
    the molecule rotation of CalcFunctions.cpp:331-339 (midpoint rule)."""
    t = (np.arange(n) + 0.5) / n
    a, b, g = 2 * np.pi * t, 2 * np.pi * t, np.pi * t
    A, B, G = np.meshgrid(a, b, g, indexing="ij")
    ca, sa, cb, sb, cg, sg = np.cos(A), np.sin(A), np.cos(B), np.sin(B), np.cos(G), np.sin(G)
    R = np.stack([np.stack([ca * cg - sa * cb * sg, -ca * sg - sa * cb * cg, sa * sb], -1),
                  np.stack([sa * cg + ca * cb * sg, -sa * sg + ca * cb * cg, -ca * sb], -1),
                  np.stack([sb * sg, sb * cg, cb], -1)], -2)
    v = (R @ np.asarray(bond)).reshape(-1, 3)
    return v.T @ v / len(v)


def test_molecules_match_reference_distribution():
    g, blocks = load("molecules")
    ref = json.loads(str(g["stats_json"]))
    bond = np.asarray(g["bond"])
    B = blocks[0]
    B.ncell = (40, 40, 40)  # a larger sample than the reference's 20^3
    el, blk, xyz = O.calc_atoms([B], seed=99)
    na, cl = xyz[el == 0], xyz[el == 1]
    b = cl - na
    assert np.abs(np.linalg.norm(b, axis=1) - np.linalg.norm(bond)).max() < 1e-12
    assert ref["bond_len_max_dev"] < 1e-12
    M = rotation_moment(bond)
    ours = b.T @ b / len(b)
    theirs = np.asarray(ref["bond_mom2"])
    s2 = float(bond @ bond)
    assert np.abs(ours - M).max() < 5 * s2 / np.sqrt(len(b))
    assert np.abs(theirs - M).max() < 5 * s2 / np.sqrt(ref["n"])
    # a uniform rotation would give s2/3 on the diagonal; this one does not
    assert np.abs(np.diag(M) - s2 / 3).max() > 10 * s2 / np.sqrt(len(b))
    # molecule displacement, sd 0.2, on the atom at the molecule origin
    disp = na - np.round(na / 6.0) * 6.0
    mr, mr2, mc2 = half_normal_stats(disp)
    sd = 0.2
    n = len(disp)
    assert abs(mr - sd * np.sqrt(2 / np.pi)) < 5 * sd * np.sqrt(1 - 2 / np.pi) / np.sqrt(n)
    assert abs(ref["mean_r"] - sd * np.sqrt(2 / np.pi)) < \
        5 * sd * np.sqrt(1 - 2 / np.pi) / np.sqrt(ref["n"])
    assert abs(mc2 - 0.5) < 5 * np.sqrt(1 / 8) / np.sqrt(n)
