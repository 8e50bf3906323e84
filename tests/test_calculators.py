"""Host logic of xansons_hip.calculators (no GPU): the calculators run on a
stand-in context whose pair sums, histograms, PDFs and 2D sums come from the
oracle, so what is checked is the host side -- bin count, weights, the FP64
combination, polarisation, Euler grid, polar q mesh and normalisation --
through the io writers against the reference executable's printed files
(tests/golden/io_files.npz), and the 2D path against or_calc_int2d, which
restates calcInt2D whole."""
import json
import os

import numpy as np
import pytest

import xs_oracle as O
from conftest import REPO
from xansons_hip import calculators as calc
from xansons_hip import io
from xansons_hip.atoms import Block


class OracleCtx:
    """The Context methods the calculators call, evaluated by the oracle."""

    def hist(self, xyz, el_count, bin_, nhist, f64=True):
        assert f64
        return O.hist_f64(np.ascontiguousarray(xyz, np.float64), el_count, bin_, nhist)

    def hist_pairsum(self, hist, bin_, q):
        return O.hist_pairsum(hist, bin_, q)

    def pdf(self, hist, el_count, bin_, pdf_type, rho0, sl, cpu_radius=True):
        return O.pdf(hist, el_count, bin_, pdf_type, rho0, sl, cpu_radius_offset=cpu_radius)

    def debye_pairsum(self, xyz, el_count, q):
        return O.debye_pairsum(np.ascontiguousarray(xyz, np.float64), el_count, q)

    def amp2d(self, xyz, el_count, qvec, rot, ff):
        return O.amp2d(np.ascontiguousarray(xyz, np.float64), el_count, qvec, rot, ff)


@pytest.fixture(scope="module")
def golden_case():
    g = np.load(os.path.join(REPO, "tests", "golden", "io_files.npz"), allow_pickle=False)
    files = json.loads(str(g["files_json"]))
    r = np.random.default_rng(404).uniform(0, 14.0, (3, 240))
    el = np.array([0] * 130 + [1] * 110, dtype=np.uint32)
    b = Block(r_atom=np.ascontiguousarray(r.T), el=el)
    e, _, xyz = O.calc_atoms([b], seed=1)
    order = np.argsort(e, kind="stable")
    counts = [int(np.sum(e == k)) for k in range(2)]
    return files, np.ascontiguousarray(xyz[order].T), counts


@pytest.mark.parametrize("pdf_type", ["RDF", "PDF", "rPDF"])
def test_pdf_and_debye_files(tmp_path, golden_case, pdf_type):
    files, xyz, counts = golden_case
    p = tmp_path / "sl.txt"
    p.write_text("Na 3.63\nCl 9.577\n")
    _, sl = io.read_sl(str(p), ["Na", "Cl"])
    t = {v: k for k, v in io.PDF_NAMES.items()}[pdf_type]
    q = calc.q_mesh(0.5, 8.0, 16)
    I, pdf, ovf = calc.calc_pdf_and_debye(OracleCtx(), xyz, counts, q, 0.02, pdf_type=t, rho0=0.08,
                                          sl=sl)
    assert ovf == 0
    name = f"io{pdf_type}"
    paths = [io.write_i(name, q, I, "neutron", str(tmp_path)),
             io.write_pdf(name, pdf[0], 0.02, "neutron", t, str(tmp_path))]
    paths += io.write_partial_pdf(name, pdf[1:], 0.02, "neutron", t, ["Na", "Cl"], str(tmp_path))
    for path in paths:
        fn = os.path.basename(path)
        assert open(path).read() == files[fn], fn


def test_2d_files_and_oracle(tmp_path, golden_case):
    files, xyz, counts = golden_case
    sl = np.array([3.63, 9.577])
    q = calc.q_mesh(0.5, 6.0, 8)
    lam = calc.default_wavelength(6.0)
    I2D, I = calc.calc_int_2d(OracleCtx(), xyz, counts, q, 6, lam, sl=sl)
    for path in (io.write_i2d("io2D", I2D, "neutron", str(tmp_path)),
                 io.write_i("io2D", q, I, "neutron", str(tmp_path))):
        fn = os.path.basename(path)
        assert open(path).read() == files[fn], fn
    # several orientations, x-ray weights and the polar factor against the
    # oracle's whole-calcInt2D restatement
    ff = np.abs(np.sin(np.outer([1.0, 2.0], q))) + 1.0
    emin, emax, en = (0.1, 0.2, 0.3), (1.0, 2.0, 0.9), (3, 2, 2)
    for conv in (0, 5, 11):
        I2D, I = calc.calc_int_2d(OracleCtx(), xyz, counts, q, 6, lam, emin, emax, en, conv,
                                  source=calc.XRAY, ff=ff, polar=True)
        O2D, O1D = O.calc_int2d(xyz, counts, q, 6, emin, emax, en, conv, lam, polar=True, ff=ff)
        assert np.max(np.abs(I2D - O2D) / np.abs(O2D)) < 1e-13
        assert np.max(np.abs(I - O1D) / np.abs(O1D)) < 1e-13


def test_debye_and_rotations_against_oracle(golden_case):
    _, xyz, counts = golden_case
    q = calc.q_mesh(0.5, 10.0, 64)
    ff = np.abs(np.cos(np.outer([0.5, 1.5], q))) + 2.0
    for polar in (False, True):
        I = calc.calc_int_debye(OracleCtx(), xyz, counts, q, source=calc.XRAY, ff=ff, polar=polar,
                                lam=calc.default_wavelength(10.0))
        S = O.debye_pairsum(xyz, counts, q)
        Io = O.intensity_1d(S, counts, q, ff=ff, polar=polar, lam=calc.default_wavelength(10.0))
        assert np.max(np.abs(I - Io) / np.abs(Io)) < 1e-14
    for conv in range(12):
        e = (0.3, -1.1, 2.4)
        assert np.max(np.abs(calc.rotation(e, conv) - O.rot_matrix(e, conv))) < 1e-15


def test_density_and_hist_size_match_oracle(golden_case):
    _, xyz, _ = golden_case
    assert calc.hist_size(xyz, 0.02) == O.hist_size(xyz, 0.02)
    assert calc.atomic_density(xyz) == O.atomic_density(xyz)
