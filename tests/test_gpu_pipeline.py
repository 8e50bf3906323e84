"""The reference's pipeline end to end on the GPU, against its printed files
(tests/golden/io_files.npz, written by the reference executable): the atoms
file through the GPU ensemble builder, the scattering-length table through
xansons_hip.io.read_sl, the calculators of xansons_hip.calculators on the
GPU, and the output files through the io writers, compared byte for byte
(the 2D files to 2e-6 of their values, see test_2d_files).

Inputs are those of tests/golden/make_io_golden.py: 240 atoms (130 Na then
130..239 Cl) drawn uniformly in a 14 A cube with numpy's default_rng(404),
neutron, the table "Na 3.63 / Cl 9.577"; DebyePDF with hist_bin 0.02, q
0.5-8 A^-1 (16 points), density 0.08 and partial PDFs, once per PDF type;
2D on q 0.5-6 A^-1 (8 points) x 6 phi with one ZXZ orientation (0, 0, 0)."""
import json
import os

import numpy as np
import pytest
try:
    import torch  # noqa: F401  (first import can take a minute on a fresh box)
except ImportError:
    torch = None

from conftest import REPO, rel_err
from xansons_hip import calculators as calc
from xansons_hip import io
from xansons_hip.atoms import Atoms, Block

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(REPO, "tests", "golden")


@pytest.fixture(scope="module")
def case(tmp_path_factory):
    g = np.load(os.path.join(GOLDEN, "io_files.npz"), allow_pickle=False)
    files = json.loads(str(g["files_json"]))
    tmp = tmp_path_factory.mktemp("pipeline")
    sl_path = tmp / "sl.txt"
    sl_path.write_text("Na 3.63\nCl 9.577\n")
    r = np.random.default_rng(404).uniform(0, 14.0, (3, 240))
    el = np.array([0] * 130 + [1] * 110, dtype=np.uint32)
    return files, str(sl_path), r, el, tmp


def built_atoms(ctx, r, el):
    """The atoms as the reference's builder leaves them (a block given as an
    atom list: one cell, the 0.5 A merge), built on the GPU."""
    a = Atoms(ctx, [Block(r_atom=np.ascontiguousarray(r.T), el=el)], 2, seed=1)
    xyz, counts = a.xyz(), [int(c) for c in a.el_count]
    a.close()
    return xyz, counts


@pytest.mark.parametrize("pdf_type", ["RDF", "PDF", "rPDF"])
def test_debye_pdf_files(ctx, case, pdf_type):
    files, sl_path, r, el, tmp = case
    names = ["Na", "Cl"]
    ids, sl = io.read_sl(sl_path, names)
    assert ids == {"Na": 0, "Cl": 1}
    xyz, counts = built_atoms(ctx, r, el)
    print(f"pipeline {pdf_type}: {sum(counts)} atoms {counts}", flush=True)
    q = calc.q_mesh(0.5, 8.0, 16)
    t = {v: k for k, v in io.PDF_NAMES.items()}[pdf_type]
    I, pdf, ovf = calc.calc_pdf_and_debye(ctx, xyz, counts, q, 0.02, scenario=calc.DEBYE_PDF,
                                          pdf_type=t, rho0=0.08, source=calc.NEUTRON, sl=sl)
    assert ovf == 0
    name = f"io{pdf_type}"
    out = str(tmp / pdf_type)
    os.makedirs(out, exist_ok=True)
    written = [io.write_i(name, q, I, "neutron", out),
               io.write_pdf(name, pdf[0], 0.02, "neutron", t, out)]
    written += io.write_partial_pdf(name, pdf[1:], 0.02, "neutron", t, names, out)
    for p in written:
        fn = os.path.basename(p)
        assert open(p).read() == files[fn], fn
    assert len(written) == 5


def test_2d_files(ctx, case):
    files, sl_path, r, el, tmp = case
    _, sl = io.read_sl(sl_path, ["Na", "Cl"])
    xyz, counts = built_atoms(ctx, r, el)
    q = calc.q_mesh(0.5, 6.0, 8)
    I2D, I = calc.calc_int_2d(ctx, xyz, counts, q, 6, calc.default_wavelength(6.0),
                              convention=0, source=calc.NEUTRON, sl=sl)
    out = str(tmp / "2d")
    os.makedirs(out, exist_ok=True)
    # the 2D sums take the FP32 hardware sine of an FP64-reduced phase (about
    # 4e-7 relative), so the printed 7 digits can differ in the last place:
    # values within 2e-6, layout (rows, columns, tabs) identical
    for p in (io.write_i2d("io2D", I2D, "neutron", out), io.write_i("io2D", q, I, "neutron", out)):
        fn = os.path.basename(p)
        got, want = open(p).read(), files[fn]
        assert [ln.count("\t") for ln in got.splitlines()] == [ln.count("\t") for ln in want.splitlines()]
        g, w = np.loadtxt(got.splitlines(), ndmin=2), np.loadtxt(want.splitlines(), ndmin=2)
        print(f"{fn}: max rel diff {np.max(np.abs(g - w) / np.abs(w)):.2e}", flush=True)
        assert rel_err(g, w)[0] < 2e-6, fn
