"""GPU pair sums of RearrangementInt (xs_cutoff_pairsum, xs_pairlist_pairsum)
against the oracle, and the whole correction on GPU sums against the
reference's own RearrangementInt output.

Tolerances: pair counts exact (the distance test is the reference's FP64
expression); sums <= 1e-12 relative to their largest magnitude (the GPU sums
each element pair in another order); the corrected intensity <= 1e-11
relative to the reference's."""
import os
import sys

import numpy as np
import pytest
try:
    import torch  # noqa: F401  (loaded before libxs_hip, as in the other GPU test modules)
except ImportError:
    torch = None

from conftest import REPO
from test_rearr_oracle import CASES, load, run

sys.path.insert(0, os.path.join(REPO, "oracle"))
import xs_oracle as O  # noqa: E402
import xansons_hip as xh  # noqa: E402
from xansons_hip import rearrangement as R  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xh.Context(0)
    yield c
    c.close()


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("case", CASES)
def test_sums_equal_oracle_on_golden_blocks(ctx, case):
    g, blocks = load(case)
    n_el = int(g["nel"])
    for B in blocks:
        if not B.rearrangement:
            continue
        r = np.ascontiguousarray(B.r_atom.T)
        r2 = R.rmax2(B.r_atom, B.neighbours, B.rcorr)
        Sg, ng = ctx.cutoff_pairsum(r, B.el, n_el, r2, g["q"])
        So, no = O.cutoff_pairsum(r, B.el, n_el, r2, g["q"])
        assert (ng == no).all() and no.sum() > 0
        assert rel(Sg, So) <= 1e-12
        rr, tt = R.neighbour_list(B.r_atom, B.el, B.neighbours, n_el)
        Lg = ctx.pairlist_pairsum(rr, tt, n_el * n_el, g["q"])
        Lo = O.pairlist_pairsum(rr, tt, n_el * n_el, g["q"])
        assert rel(Lg, Lo) <= 1e-12


@pytest.mark.parametrize("case", CASES)
def test_rearrangement_on_gpu_equals_reference(ctx, case):
    g, blocks = load(case)
    got = run(g, blocks, ctx.cutoff_pairsum, ctx.pairlist_pairsum)
    assert rel(got, g["rearr_I"]) <= 1e-11


def test_large_block_equals_oracle(ctx):
    """6,000 atoms, 3 elements, cut-offs of 2-4 A: many tiles and chunks."""
    rng = np.random.default_rng(21)
    n = 6000
    r = rng.uniform(0, 45.0, (3, n))
    el = np.sort(rng.integers(0, 3, n)).astype(np.uint32)
    r2 = rng.uniform(4.0, 16.0, n)
    q = np.linspace(0.3, 12.0, 300)
    Sg, ng = ctx.cutoff_pairsum(r, el, 3, r2, q)
    So, no = O.cutoff_pairsum(r, el, 3, r2, q)
    assert (ng == no).all() and no.sum() > 10000
    assert rel(Sg, So) <= 1e-12


def test_rejects_bad_arguments(ctx):
    with pytest.raises(xh.XSError):
        ctx.cutoff_pairsum(np.zeros((3, 2)), [0, 3], 2, [1.0, 1.0], [1.0])
    with pytest.raises(xh.XSError):
        ctx.pairlist_pairsum([1.0], [4], 4, [1.0])
