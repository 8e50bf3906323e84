"""GPU ensemble builder (xs_atoms_build) against the reference's own atoms
(golden fixtures from tests/golden/make_atoms_golden.py) and against the
oracle (or_calc_atoms) with the same Philox seed.

Tolerances: element counts, order and every deterministic position are
compared bit for bit.  Positions that pass through sin / cos / log (atom
and molecule displacements, molecule rotations) may differ from the
oracle's libm by an ulp of those functions: |dx| <= 1e-12 A."""
import json
import os
import sys

import numpy as np
import pytest
try:
    import torch  # loaded before libxs_hip, as in the other GPU test modules
except ImportError:
    torch = None

from conftest import REPO

sys.path.insert(0, os.path.join(REPO, "oracle"))
import xs_oracle as O  # noqa: E402
import xansons_hip as xh  # noqa: E402
from xansons_hip.atoms import Atoms, Block  # noqa: E402

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(REPO, "tests", "golden")
DETERMINISTIC = ["rocksalt", "hex_copies", "shared_sites", "single_cell"]


@pytest.fixture(scope="module")
def ctx():
    c = xh.Context(0)
    yield c
    c.close()


def load(case):
    g = np.load(os.path.join(GOLDEN, f"atoms_{case}.npz"))
    return g, [Block.from_dict(d) for d in json.loads(str(g["blocks_json"]))]


def oracle_element_major(blocks, n_el, seed):
    el, blk, xyz = O.calc_atoms(blocks, seed)
    order = np.argsort(el, kind="stable")
    per_block = np.zeros((len(blocks), n_el), np.int64)
    np.add.at(per_block, (blk, el), 1)
    return xyz[order], np.bincount(el, minlength=n_el), per_block


def gpu(ctx, blocks, n_el, seed):
    a = Atoms(ctx, blocks, n_el, seed)
    try:
        return a.xyz().T.copy(), a.el_count.astype(np.int64), a.block_el_count.astype(np.int64)
    finally:
        a.close()


@pytest.mark.parametrize("case", DETERMINISTIC)
def test_gpu_equals_reference_atoms(ctx, case):
    g, blocks = load(case)
    n_el = int(g["nel"])
    xyz, counts, per_block = gpu(ctx, blocks, n_el, seed=2024)
    assert (counts == g["el_count"]).all()
    assert (per_block == g["block_el_count"]).all()
    assert xyz.shape == g["atoms"].shape
    assert (xyz.view(np.uint64) == g["atoms"].view(np.uint64)).all()


def random_block(rng, n_el, natom, ncell, nsymm=1, ncopy=1, occ=True, dev=True, mol=False,
                 rcut=0.0, rcut_copies=0.0, scale=5.0):
    """A block exercising every branch: partial and shared occupancies,
    symmetry operations (with special positions), copies, cut-offs."""
    cell = np.diag([scale, scale * 1.1, scale * 0.9]) + rng.uniform(-0.4, 0.4, (3, 3))
    fr = rng.uniform(0, 1, (natom, 3))
    fr[1::7] = fr[0::7][: len(fr[1::7])] + rng.uniform(-0.03, 0.03, (len(fr[1::7]), 3))  # shared
    el = np.sort(rng.integers(0, n_el, natom)).astype(np.uint32)
    r = fr @ cell
    P = [np.eye(3)]
    T = [np.zeros(3)]
    for _ in range(nsymm - 1):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        P.append(np.round(q))  # signed permutation-like matrices, so images coincide
        T.append(rng.integers(0, 2, 3) * 0.5 @ cell)
    rc = rng.uniform(-20, 20, (ncopy, 3))
    cf = []
    for _ in range(ncopy):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        cf.append(q)
    return Block(r_atom=r, el=el,
                 occ=rng.uniform(0.2, 0.7, natom) if occ else np.ones(natom),
                 dev=rng.uniform(0.0, 0.1, natom) * (rng.uniform(size=natom) > 0.3) if dev else
                 np.zeros(natom),
                 p_symm=np.array(P), r_symm=np.array(T), cell=cell, ncell=ncell,
                 r_copy=rc, cf=np.array(cf), r_mean=r.mean(0) + 0.5 * np.array(ncell) @ cell,
                 rcut=rcut, rcut_copies=rcut_copies, dev_mol=0.08 if mol else 0.0,
                 mol_rotation=mol)


def compare_with_oracle(ctx, blocks, n_el, seed, exact):
    want, wc, wb = oracle_element_major(blocks, n_el, seed)
    got, gc, gb = gpu(ctx, blocks, n_el, seed)
    assert (gc == wc).all() and (gb == wb).all(), (gc, wc)
    assert got.shape == want.shape
    if exact:
        assert (got.view(np.uint64) == want.view(np.uint64)).all()
    else:
        assert np.abs(got - want).max() <= 1e-12
    return len(got)


def test_gpu_equals_oracle_random_blocks(ctx):
    rng = np.random.default_rng(11)
    blocks = [random_block(rng, 3, 40, (4, 3, 5), nsymm=4, ncopy=3, rcut=9.0, rcut_copies=30.0),
              random_block(rng, 3, 25, (6, 6, 2), nsymm=2, mol=True),
              random_block(rng, 2, 60, (1, 1, 1), occ=False, dev=False)]
    for seed in (1, 2 ** 63 + 5):
        assert compare_with_oracle(ctx, blocks, 3, seed, exact=False) > 1000


def test_gpu_equals_oracle_occupancy_only_bit_exact(ctx):
    # occupancy draws and merging only (no transcendental functions): exact
    rng = np.random.default_rng(12)
    blocks = [random_block(rng, 2, 50, (5, 4, 3), nsymm=3, ncopy=2, dev=False, rcut=12.0),
              random_block(rng, 2, 30, (3, 3, 3), nsymm=2, dev=False)]
    assert compare_with_oracle(ctx, blocks, 2, 77, exact=True) > 500


@pytest.mark.parametrize("case", ["random_occ", "molecules"])
def test_gpu_equals_oracle_golden_random_inputs(ctx, case):
    g, blocks = load(case)
    compare_with_oracle(ctx, blocks, 2, 31337, exact=False)


def test_gpu_single_cell_large(ctx):
    """A finished 12,000-atom structure given as one cell, with 1,200
    planted near-duplicates (the O(N^2) case of the reference)."""
    rng = np.random.default_rng(13)
    n = 12000
    r = rng.uniform(0, 60.0, (n, 3))
    dup = rng.choice(n, 1200, replace=False)
    r2 = r[dup] + rng.normal(size=(1200, 3)) * 0.15
    el = np.sort(rng.integers(0, 2, n + 1200)).astype(np.uint32)
    B = Block(r_atom=np.concatenate([r, r2]), el=el)
    m = compare_with_oracle(ctx, [B], 2, 5, exact=True)
    assert m < n + 1200


def test_gpu_periodic_images_merge(ctx):
    # sites near opposite faces of an oblique cell coincide under a
    # translation by one cell (125 translations tested per pair)
    rng = np.random.default_rng(14)
    cell = np.array([[6.0, 0, 0], [1.0, 5.0, 0], [0.5, 0.3, 4.0]])
    fr = rng.uniform(0, 1, (400, 3))
    fr[:40] = [0.005, 0.5, 0.5] + rng.uniform(0, 0.01, (40, 3))
    fr[40:80] = [0.995, 0.5, 0.5] + rng.uniform(-0.01, 0, (40, 3))
    el = np.sort(rng.integers(0, 2, 400)).astype(np.uint32)
    B = Block(r_atom=fr @ cell, el=el, occ=rng.uniform(0.3, 1.0, 400), cell=cell, ncell=(3, 3, 3))
    compare_with_oracle(ctx, [B], 2, 8, exact=True)


def test_gpu_device_copy(ctx):
    g, blocks = load("rocksalt")
    a = Atoms(ctx, blocks, 2, 0)
    try:
        t = torch.empty(3 * a.n, dtype=torch.float64, device="cuda")
        a.copy_dev(t.data_ptr())
        ctx.synchronize()
        assert (t.cpu().numpy().reshape(3, -1).T == g["atoms"]).all()
    finally:
        a.close()


def test_gpu_build_rejects_bad_input(ctx):
    B = Block(r_atom=np.zeros((2, 3)), el=[0, 5])
    with pytest.raises(xh.XSError) as e:
        Atoms(ctx, [B], 2, 0)
    assert e.value.code == -1
    B = Block(r_atom=[[0, 0, np.nan]], el=[0])
    with pytest.raises(xh.XSError):
        Atoms(ctx, [B], 1, 0)
