"""The reference's calculators for Python callers, on the GPU.

calc_int_debye, calc_pdf_and_debye and calc_int_2d take what the reference's
calcIntDebye, calcPDFandDebye and calcInt2D take (main.cpp:65-68,
CalcFunctions.cpp:547, 912, 969): element-major atoms, the q mesh, the x-ray
form factors FF[el][iq] or neutron scattering lengths SL[el], and the
settings of the <Calculation> element, and return what those functions
allocate and fill: I(q), the PDFs (total first, then the partials) and the
2D pattern.  The pair sums, histograms and PDFs run on the GPU through
libxs_hip; the combination stays on the host in FP64 in the operation order
of the C++ drop-in (dropin/CalcFunctionsHIP.cpp), which is that of the CPU
path (CalcFunctions.cpp:592-626, 872-906, 999-1075), so results equal the
reference's to the pair sums' rounding.  With the table readers and writers
of xansons_hip.io this is the reference's pipeline: read FF/SL, calculate,
print.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence, Tuple

import numpy as np

from . import Context

XRAY, NEUTRON = "xray", "neutron"
# scenario values of typedefs.h:49-52
DEBYE, DEBYE_HIST, PDF_ONLY, DEBYE_PDF, TWO_D = 0, 1, 2, 3, 4
PI = 3.141592653589793238462643383279  # typedefs.h PI


def q_mesh(qmin: float, qmax: float, n: int) -> np.ndarray:
    """The reference's q mesh (ReadXML.cpp:220-222): qmin + iq (qmax - qmin)/N,
    iq < N (qmax itself excluded)."""
    dq = (qmax - qmin) / n
    return np.array([qmin + iq * dq for iq in range(n)])


def default_wavelength(qmax: float) -> float:
    """The reference's default wavelength 4 pi / qmax of the <q> element
    (ReadXML.cpp:219)."""
    return 4. * PI / qmax


def _lam(polar: bool, lam: Optional[float]) -> Optional[float]:
    if polar and lam is None:
        raise ValueError("the polarisation factor needs the wavelength (default_wavelength(qmax))")
    return lam


def _weights(source: str, ff, sl, n_el: int, nq: int) -> np.ndarray:
    """weight(el, iq): FF[el][iq] for x-rays, SL[el] for neutrons."""
    if source == XRAY:
        w = np.ascontiguousarray(ff, dtype=np.float64).reshape(n_el, nq)
    else:
        w = np.repeat(np.ascontiguousarray(sl, dtype=np.float64).reshape(n_el, 1), nq, axis=1)
    return w


def polar_1d(I: np.ndarray, q: np.ndarray, lam: float) -> np.ndarray:
    """PolarFactor1D (CalcFunctions.cpp:451-458), in place."""
    st = q * (lam * 0.25 / PI)
    c2t = 1. - 2. * (st * st)
    I *= 0.5 * (1. + c2t * c2t)
    return I


def hist_size(xyz: np.ndarray, hist_bin: float) -> int:
    """GetHistSize (CalcFunctions.cpp:224-243): the bounding-box diagonal in
    bins, plus one."""
    d = np.asarray(xyz, np.float64).max(axis=1) - np.asarray(xyz, np.float64).min(axis=1)
    return int(math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / hist_bin) + 1


def atomic_density(xyz: np.ndarray) -> float:
    """GetAtomicDensity (CalcFunctions.cpp:196-223): atoms within a third of
    the ensemble size around the centroid, over that sphere's volume.  The
    centroid is summed atom by atom, as the reference does."""
    x = np.asarray(xyz, np.float64)
    n = x.shape[1]
    c = np.cumsum(x, axis=1)[:, -1] / n
    d2 = (c[0] - x[0]) ** 2 + (c[1] - x[1]) ** 2 + (c[2] - x[2]) ** 2
    size = 2. * math.sqrt(float(d2.max()))
    rmax2 = (0.33333 * size) ** 2
    count = int(np.count_nonzero(d2 < rmax2))
    rmax = math.sqrt(rmax2)
    return count / (4. / 3. * PI * rmax * rmax * rmax)


def _combine(S: np.ndarray, w: np.ndarray, el_count: Sequence[int], ntot: int,
             twice_first: bool) -> np.ndarray:
    """I(q) = (sum_p 2 w_a w_b S_p + sum_e N_e w_e^2) / Ntot in the order of
    the drop-in: calcPDFandDebye forms 2 S w_a w_b, calcIntDebye S w_a w_b 2."""
    n_el = len(el_count)
    I = np.zeros(S.shape[1])
    p = 0
    for a in range(n_el):
        for b in range(a, n_el):
            if twice_first:
                I += 2. * S[p] * w[a] * w[b]
            else:
                I += S[p] * w[a] * w[b] * 2.
            p += 1
    for e in range(n_el):
        I += el_count[e] * (w[e] * w[e])
    return I / ntot


def calc_int_debye(ctx: Context, xyz: np.ndarray, el_count: Sequence[int], q: np.ndarray,
                   source: str = NEUTRON, ff=None, sl=None, polar: bool = False,
                   lam: Optional[float] = None) -> np.ndarray:
    """calcIntDebye (CalcFunctions.cpp:547-633): direct Debye I(q)."""
    q = np.ascontiguousarray(q, np.float64)
    S = ctx.debye_pairsum(xyz, el_count, q)
    w = _weights(source, ff, sl, len(el_count), q.size)
    I = _combine(S, w, el_count, int(sum(el_count)), twice_first=False)
    if polar:
        polar_1d(I, q, _lam(polar, lam))
    return I


def calc_pdf_and_debye(ctx: Context, xyz: np.ndarray, el_count: Sequence[int], q: np.ndarray,
                       hist_bin: float, scenario: int = DEBYE_PDF, pdf_type: int = 0,
                       rho0: Optional[float] = None, source: str = NEUTRON, ff=None, sl=None,
                       sl_pdf=None, polar: bool = False, lam: Optional[float] = None,
                       nhist: Optional[int] = None, cpu_radius: bool = True
                       ) -> Tuple[Optional[np.ndarray], Optional[np.ndarray], int]:
    """calcPDFandDebye (CalcFunctions.cpp:912-967): the FP64-binned histogram
    (counts equal calcHist's), then I(q) from it for Debye_hist / DebyePDF
    and the PDFs ((1 + ntypes) x nhist, total first) for PDFonly / DebyePDF.
    sl_pdf: the PDF weights (the scattering lengths for neutrons; for x-rays
    the table the reference reads for them).  rho0 defaults to the
    reference's GetAtomicDensity.  Returns (I or None, PDF or None,
    overflow count)."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    q = np.ascontiguousarray(q, np.float64)
    nh = nhist if nhist is not None else hist_size(xyz, hist_bin)
    hist, ovf = ctx.hist(xyz, el_count, hist_bin, nh, f64=True)
    I = pdf = None
    if scenario > DEBYE_HIST:
        rho = rho0 if rho0 is not None else atomic_density(xyz)
        pdf = ctx.pdf(hist, el_count, hist_bin, pdf_type, rho,
                      sl_pdf if sl_pdf is not None else sl, cpu_radius=cpu_radius)
    if scenario in (DEBYE_HIST, DEBYE_PDF):
        S = ctx.hist_pairsum(hist, hist_bin, q)
        w = _weights(source, ff, sl, len(el_count), q.size)
        I = _combine(S, w, el_count, int(sum(el_count)), twice_first=True)
        if polar:
            polar_1d(I, q, _lam(polar, lam))
    return I, pdf, ovf


def rotation(euler: Sequence[float], convention: int) -> np.ndarray:
    """calcRotMatrix (CalcFunctions.cpp:46-115) as the product of the three
    elementary rotations of an intrinsic convention (typedefs.h:59-70
    numbering); row-major 3 x 3."""
    axes = ((0, 2, 0), (0, 1, 0), (1, 0, 1), (1, 2, 1), (2, 1, 2), (2, 0, 2),
            (0, 2, 1), (0, 1, 2), (1, 0, 2), (1, 2, 0), (2, 1, 0), (2, 0, 1))
    acc = [[1., 0., 0.], [0., 1., 0.], [0., 0., 1.]]
    for s in range(3):
        ax = axes[convention][s]
        u, v = (ax + 1) % 3, (ax + 2) % 3
        c, sn = math.cos(euler[s]), math.sin(euler[s])
        M = [[0., 0., 0.], [0., 0., 0.], [0., 0., 0.]]
        M[ax][ax] = 1.
        M[u][u], M[u][v], M[v][u], M[v][v] = c, -sn, sn, c
        acc = [[acc[i][0] * M[0][j] + acc[i][1] * M[1][j] + acc[i][2] * M[2][j] for j in range(3)]
               for i in range(3)]
    return np.array(acc)


def euler_grid(emin: Sequence[float], emax: Sequence[float], n: Sequence[int],
               convention: int) -> np.ndarray:
    """Rotations of the Euler midpoint grid (CalcFunctions.cpp:1007-1017):
    step 0 on an axis with fewer than two points."""
    d = [0. if n[k] < 2 else (emax[k] - emin[k]) / n[k] for k in range(3)]
    rots = []
    for ia in range(n[0]):
        for ib in range(n[1]):
            for ig in range(n[2]):
                e = (emin[0] + (ia + 0.5) * d[0], emin[1] + (ib + 0.5) * d[1],
                     emin[2] + (ig + 0.5) * d[2])
                rots.append(rotation(e, convention))
    return np.array(rots).reshape(-1, 9)


def calc_int_2d(ctx: Context, xyz: np.ndarray, el_count: Sequence[int], q: np.ndarray, nfi: int,
                lam: float, euler_min: Sequence[float] = (0., 0., 0.),
                euler_max: Sequence[float] = (0., 0., 0.), euler_n: Sequence[int] = (1, 1, 1),
                convention: int = 0, source: str = NEUTRON, ff=None, sl=None, polar: bool = False
                ) -> Tuple[np.ndarray, np.ndarray]:
    """calcInt2D (CalcFunctions.cpp:969-1091): the polar (q, phi) mesh with
    cos(theta) = 1 - sin^2(theta), q vectors rotated by R^T for every
    orientation of the Euler grid (angles in radians), |A|^2 summed, divided
    by Ntot Nrot, polar factor, then I(q) as the phi average.  lam: the
    wavelength (default_wavelength(qmax) unless the input gives one).
    Returns (I2D (nq, nfi), I (nq))."""
    q = np.ascontiguousarray(q, np.float64)
    nq = q.size
    w1 = _weights(source, ff, sl, len(el_count), nq)
    dfi = 2. * PI / nfi
    cphi = np.array([math.cos(i * dfi) for i in range(nfi)])
    sphi = np.array([math.sin(i * dfi) for i in range(nfi)])
    st = q * (lam * 0.25 / PI)
    ct = 1. - st * st
    qvec = np.empty((nq, nfi, 3))
    qvec[:, :, 0] = (ct[:, None] * cphi[None, :]) * q[:, None]
    qvec[:, :, 1] = (ct[:, None] * sphi[None, :]) * q[:, None]
    qvec[:, :, 2] = (-st * q)[:, None]
    w = np.repeat(w1[:, :, None], nfi, axis=2).reshape(len(el_count), nq * nfi)
    rot = euler_grid(euler_min, euler_max, euler_n, convention)
    acc = ctx.amp2d(xyz, el_count, qvec.reshape(-1, 3), rot, w).reshape(nq, nfi)
    norm = 1. / (int(sum(el_count)) * euler_n[0] * euler_n[1] * euler_n[2])
    I2D = acc * norm
    if polar:
        c2t = 1. - 2. * (st * st)
        I2D = I2D * (0.5 * (1. + c2t * c2t))[:, None]
    I = np.cumsum(I2D, axis=1)[:, -1] / nfi
    return I2D, I
