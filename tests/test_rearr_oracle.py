"""RearrangementInt (CalcFunctions.cpp:469-545): the oracle's pair sums
(or_cutoff_pairsum, or_pairlist_pairsum) through the package's host
combination (xansons_hip.rearrangement) against the reference's own
RearrangementInt output on I0[k] = 1 + k/64 (tests/golden/
make_rearr_golden.py): neutron with the sorted-index neighbour quirk, x-ray
with polarisation and partial intensities over two blocks, a 500-atom
three-element cluster."""
import json
import os
import sys

import numpy as np
import pytest

from conftest import REPO

sys.path.insert(0, os.path.join(REPO, "oracle"))
import xs_oracle as O  # noqa: E402
from xansons_hip import rearrangement as R  # noqa: E402

GOLDEN = os.path.join(REPO, "tests", "golden")
CASES = ["nacl", "xray_partial", "cluster"]


def hexa(v):
    return np.array([float.fromhex(x) if isinstance(x, str) else float(x) for x in v])


def load(case):
    g = np.load(os.path.join(GOLDEN, f"rearr_{case}.npz"))
    blocks = []
    for d in json.loads(str(g["blocks_json"])):
        nat = int(d["natom"])
        rb = "neighbours" in d
        nbs = [np.array([[float(e)] + [float.fromhex(c) for c in xyz] for e, *xyz in lst])
               .reshape(-1, 4) for lst in d["neighbours"]] if rb else [np.zeros((0, 4))] * nat
        blocks.append(R.RearrBlock(
            r_atom=hexa(d["r_atom"]).reshape(nat, 3), el=np.asarray(d["el"], np.uint32),
            neighbours=nbs, rcorr=float.fromhex(d["rcorr"]) if rb else 0.0,
            coeff=int(d["ncopy"]) * int(np.prod(d["ncell"])) * int(d["nsymm"]), rearrangement=rb))
    return g, blocks


def inputs(g):
    n_el = int(g["nel"])
    src = "xray" if int(g["source"]) == 1 else "neutron"
    I0 = 1.0 + np.arange(g["rearr_I"].size) / 64.0
    return n_el, src, I0


def run(g, blocks, cutoff_sums, list_sums):
    n_el, src, I0 = inputs(g)
    return R.rearrangement_intensity(I0, g["q"], blocks, n_el, int(g["el_count"].sum()), src,
                                     ff=g["ff"] if src == "xray" else None,
                                     sl=g["sl"] if src == "neutron" else None,
                                     partial=bool(g["partial"]), polar=bool(g["polar"]),
                                     lam=float(g["lam"]), cutoff_sums=cutoff_sums,
                                     list_sums=list_sums)


@pytest.mark.parametrize("case", CASES)
def test_oracle_matches_reference_rearrangement(case):
    g, blocks = load(case)
    got = run(g, blocks, O.cutoff_pairsum, O.pairlist_pairsum)
    want = g["rearr_I"]
    # only the neighbour terms are summed in another order than the reference
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_cutoff_pairs_are_the_references():
    """A planted change of the pair test (<= instead of <) or of the
    self-distance threshold moves the result away from the reference's."""
    g, blocks = load("cluster")
    want = g["rearr_I"]

    def loose(xyz, el, n_el, r2, q):
        return O.cutoff_pairsum(xyz, el, n_el, np.nextafter(r2, np.inf) * (1 + 1e-3), q)
    bad = run(g, blocks, loose, O.pairlist_pairsum)
    assert np.abs(bad - want).max() > 1e-9 * np.abs(want).max()


def test_partial_layout_and_polar():
    g, blocks = load("xray_partial")
    nq = g["q"].size
    got = run(g, blocks, O.cutoff_pairsum, O.pairlist_pairsum)
    # RearrangementInt rewrites the block-diagonal parts only (i0qST,
    # CalcFunctions.cpp:535): (0,0) at nq and (1,1) at 3 nq; the (0,1) part
    # at 2 nq keeps its input
    I0 = 1.0 + np.arange(got.size) / 64.0
    assert np.array_equal(got[2 * nq:3 * nq], I0[2 * nq:3 * nq])
    assert not np.array_equal(got[nq:2 * nq], I0[nq:2 * nq])
