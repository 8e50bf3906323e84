"""The reference's table readers and output writers, byte-compatible.

Readers: scattering lengths (readNeuData, IO.cpp:296-319) and form factors
with their linear interpolation onto the q mesh (loadFF, IO.cpp:320-377).
Writers: 1D / partial 1D / 2D intensities, PDFs and partial PDFs, atoms
(IO.cpp:51-167, 273-295), with the reference's file names, columns,
``%.6e`` numbers (ostream with ios::scientific, precision 6) and tabs, so
that scripts reading the reference's files read these unchanged.
"""
from __future__ import annotations

import os
from typing import Dict, List, Sequence, Tuple

import numpy as np

PDF_NAMES = {0: "RDF", 1: "PDF", 2: "rPDF"}  # typedefs.h:54-56


def _split(line: str) -> Tuple[str, str]:
    """findSpace (IO.cpp:31-39): the element name ends at the first space or
    tab (position 0 if the line starts with a space and has a tab later)."""
    ps, pt = line.find(" "), line.find("\t")
    pos = 0
    if ps > 0 and pt > 0:
        pos = min(ps, pt)
    if pt < 0:
        pos = ps
    if ps < 0:
        pos = pt
    if pos < 0:  # no separator: substr(0, npos) and substr(0)
        return line, line
    return line[:pos], line[pos + 1:]


def _numbers(text: str) -> List[float]:
    """`while (stream >> value)`: leading numbers of a whitespace-separated list."""
    out = []
    for tok in text.split():
        try:
            out.append(float(tok))
        except ValueError:
            break
    return out


def _lines(path: str):
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if line and line[0] != "#":
                yield line


def read_sl(path: str, names: Sequence[str]) -> Tuple[Dict[str, int], np.ndarray]:
    """readNeuData: element index = order of the table rows naming elements
    of ``names``; returns ({name: index}, scattering lengths)."""
    ids, sl = {}, []
    for line in _lines(path):
        el, rest = _split(line)
        if el in names:
            ids[el] = len(sl)
            v = _numbers(rest)
            sl.append(v[0] if v else 0.0)
    return ids, np.array(sl)


def load_ff(path: str, names: Sequence[str], q) -> Tuple[Dict[str, int], np.ndarray]:
    """loadFF: form factors of the table's rows naming elements of ``names``
    (index = row order) linearly interpolated onto q with the reference's
    arithmetic; the search for the bracketing table interval only moves
    forward.  q outside the table is left NaN (the reference leaves those
    values uninitialised)."""
    q = np.asarray(q, np.float64)
    ids, qff, ff = {}, [], []
    for line in _lines(path):
        el, rest = _split(line)
        if el == "q":
            qff.extend(_numbers(rest))
        elif el in names:
            ids[el] = len(ff)
            ff.append(_numbers(rest))
    if len(ff) < len(names):
        raise ValueError(f"{path}: not enough form factors")
    out = np.full((len(ff), q.size), np.nan)
    start = 0
    for iq, qq in enumerate(q.tolist()):
        for k in range(start, len(qff) - 1):
            if qq >= qff[k] and qq <= qff[k + 1]:
                dq = qff[k] - qff[k + 1]
                for i, row in enumerate(ff):
                    a = (row[k] - row[k + 1]) / dq
                    b = (row[k + 1] * qff[k] - row[k] * qff[k + 1]) / dq
                    out[i, iq] = a * qq + b
                start = k
                break
    return ids, out


def _src(source: str) -> str:
    return "xray" if source == "xray" else "neut"


def _e(v: float) -> str:
    return "%.6e" % v


def write_i(name: str, q, I, source: str, directory: str = ".") -> str:
    """printI (IO.cpp:123-138): <name>_{xray,neut}_1D.txt, 'q<TAB>I' rows."""
    path = os.path.join(directory, f"{name}_{_src(source)}_1D.txt")
    with open(path, "w") as f:
        f.write("".join(f"{_e(a)}\t{_e(b)}\n" for a, b in zip(np.asarray(q).tolist(),
                                                                np.asarray(I).tolist())))
    return path


def write_partial_i(name: str, q, I_parts, source: str, nblocks: int, directory: str = ".") -> List[str]:
    """printPartialI (IO.cpp:139-161): one file per block pair iB <= jB,
    <name>_{xray,neut}_<iB>-<jB>_1D.txt; I_parts rows in that order."""
    paths, k = [], 0
    parts = np.asarray(I_parts).reshape(-1, np.asarray(q).size)
    for iB in range(nblocks):
        for jB in range(iB, nblocks):
            p = os.path.join(directory, f"{name}_{_src(source)}_{iB}-{jB}_1D.txt")
            with open(p, "w") as f:
                f.write("".join(f"{_e(a)}\t{_e(b)}\n"
                                for a, b in zip(np.asarray(q).tolist(), parts[k].tolist())))
            paths.append(p)
            k += 1
    return paths


def write_i2d(name: str, I2D, source: str, directory: str = ".") -> str:
    """printI2 (IO.cpp:162-181): rows q, columns phi, each value followed by a tab."""
    path = os.path.join(directory, f"{name}_{_src(source)}_2D.txt")
    with open(path, "w") as f:
        for row in np.asarray(I2D).tolist():
            f.write("".join(f"{_e(v)}\t" for v in row) + "\n")
    return path


def _pdf_rows(values, bin_: float, absolute: bool) -> str:
    v = np.asarray(values).tolist()
    last = len(v) - 1
    while last > 0:  # trailing values at or below 1e-10 are not printed
        if (abs(v[last]) if absolute else v[last]) > 1.e-10:
            break
        last -= 1
    return "".join(f"{_e((i + 0.5) * bin_)}\t{_e(v[i])}\n" for i in range(last + 1))


def write_pdf(name: str, pdf, bin_: float, source: str, pdf_type: int, directory: str = ".") -> str:
    """printPDF (IO.cpp:51-73): <name>_{xray,neut}_{RDF,PDF,rPDF}.txt,
    'r<TAB>value' rows, r = (i + 1/2) bin, trailing |value| <= 1e-10 cut."""
    path = os.path.join(directory, f"{name}_{_src(source)}_{PDF_NAMES[int(pdf_type)]}.txt")
    with open(path, "w") as f:
        f.write(_pdf_rows(pdf, bin_, True))
    return path


def write_partial_pdf(name: str, parts, bin_: float, source: str, pdf_type: int,
                      names: Sequence[str], directory: str = ".") -> List[str]:
    """printPartialPDF (IO.cpp:74-106): one file per pair type (a <= b) in
    element-index order, labelled with the a-th and b-th names of the
    alphabetically sorted element map (as the reference's std::map iterator
    does), trailing values <= 1e-10 (signed) cut."""
    nel = len(names)
    labels = sorted(names)
    rows = np.asarray(parts).reshape(nel * (nel + 1) // 2, -1)
    paths, k = [], 0
    for a in range(nel):
        for b in range(a, nel):
            p = os.path.join(directory, f"{name}_{_src(source)}_partial_{PDF_NAMES[int(pdf_type)]}_"
                                        f"{labels[a]}_{labels[b]}.txt")
            with open(p, "w") as f:
                f.write(_pdf_rows(rows[k], bin_, False))
            paths.append(p)
            k += 1
    return paths


def write_atoms(name: str, xyz, el_count: Sequence[int], names: Sequence[str],
                directory: str = ".") -> str:
    """printAtoms (IO.cpp:273-295): <name>_atoms.xyz, the atom count, an
    empty line, then 'name x  y  z' rows element by element; the el-th
    element is labelled with the el-th name of the alphabetically sorted
    element map, as the reference does."""
    xyz = np.asarray(xyz, np.float64)
    labels = sorted(names)
    path = os.path.join(directory, f"{name}_atoms.xyz")
    n = int(sum(el_count))
    out = [f"{n}\n\n"]
    o = 0
    for e, c in enumerate(el_count):
        lab = labels[e]
        for j in range(o, o + int(c)):
            out.append(f"{lab} {_e(xyz[0, j])}  {_e(xyz[1, j])}  {_e(xyz[2, j])}\n")
        o += int(c)
    with open(path, "w") as f:
        f.write("".join(out))
    return path
