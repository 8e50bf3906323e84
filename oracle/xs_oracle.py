"""ctypes binding of liboracle.so, the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; it is the checker, never the measured or shipped path.
Each wrapper mirrors one C function of xs_oracle.c; the reference file:line
each follows is given in xs_oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_hist_size.restype = C.c_uint
        L.or_hist_size.argtypes = [f64p, C.c_size_t, C.c_double]
        L.or_atomic_density.restype = C.c_double
        L.or_atomic_density.argtypes = [f64p, C.c_size_t]
        L.or_hist_f64.argtypes = [f64p, u32p, C.c_uint, C.c_double, C.c_uint, u64p,
                                  C.POINTER(C.c_uint64), C.c_int]
        L.or_hist_f32.argtypes = [f32p, u32p, C.c_uint, C.c_float, C.c_uint, u64p,
                                  C.POINTER(C.c_uint64), C.c_int]
        L.or_debye_pairsum.argtypes = [f64p, u32p, C.c_uint, f64p, C.c_uint, f64p, C.c_int]
        L.or_partial_debye_pairsum.argtypes = [f64p, u32p, C.c_uint, u32p, C.c_uint, f64p,
                                               C.c_uint, f64p, C.c_int]
        L.or_hist_pairsum.argtypes = [u64p, C.c_uint, C.c_uint, C.c_double, f64p, C.c_uint,
                                      f64p, C.c_int]
        L.or_intensity_1d.argtypes = [f64p, C.c_uint, u32p, C.c_void_p, C.c_void_p, f64p,
                                      C.c_uint, C.c_int, C.c_double, f64p]
        L.or_partial_intensity_1d.argtypes = [f64p, C.c_uint, u32p, C.c_uint, C.c_void_p,
                                              C.c_void_p, f64p, C.c_uint, C.c_int, C.c_double,
                                              C.c_uint, f64p]
        L.or_polar_1d.argtypes = [f64p, f64p, C.c_uint, C.c_double]
        L.or_pdf.argtypes = [u64p, C.c_uint, u32p, C.c_uint, C.c_double, C.c_int, C.c_double,
                             f64p, C.c_int, f64p]
        L.or_rot_matrix.argtypes = [f64p, C.c_int, f64p]
        L.or_amp2d.argtypes = [f64p, u32p, C.c_uint, f64p, C.c_size_t, f64p, C.c_uint, f64p,
                               f64p, C.c_int]
        L.or_calc_int2d.argtypes = [f64p, u32p, C.c_uint, f64p, C.c_uint, C.c_uint, f64p, f64p,
                                    u32p, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p,
                                    f64p, f64p, C.c_int]
        L.or_philox4x64.argtypes = [u64p, u64p, u64p]
        L.or_cutoff_pairsum.argtypes = [f64p, C.c_size_t, u32p, C.c_uint, f64p, f64p, C.c_uint,
                                        f64p, u64p]
        L.or_pairlist_pairsum.argtypes = [f64p, u32p, C.c_size_t, C.c_uint, f64p, C.c_uint, f64p]
        L.or_calc_atoms.restype = C.c_longlong
        L.or_calc_atoms.argtypes = [C.c_void_p, C.c_uint, C.c_uint64, C.c_longlong, u32p, u32p,
                                    f64p]
        _LIB = L
    return _LIB


def _threads(n):
    return int(n) if n else min(16, os.cpu_count() or 1)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _cu32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def ntypes(n_el: int) -> int:
    return n_el * (n_el + 1) // 2


def hist_size(xyz64, bin_: float) -> int:
    xyz64 = _c64(xyz64)
    return int(lib().or_hist_size(xyz64, xyz64.shape[-1], bin_))


def atomic_density(xyz64) -> float:
    xyz64 = _c64(xyz64)
    return float(lib().or_atomic_density(xyz64, xyz64.shape[-1]))


def hist_f64(xyz64, el_count, bin_, nhist, threads=None):
    xyz64, ec = _c64(xyz64), _cu32(el_count)
    h = np.zeros(ntypes(len(ec)) * nhist, dtype=np.uint64)
    ov = C.c_uint64(0)
    lib().or_hist_f64(xyz64, ec, len(ec), float(bin_), int(nhist), h, C.byref(ov), _threads(threads))
    return h.reshape(-1, nhist), int(ov.value)


def hist_f32(xyz32, el_count, bin_, nhist, threads=None):
    xyz32, ec = np.ascontiguousarray(xyz32, dtype=np.float32), _cu32(el_count)
    h = np.zeros(ntypes(len(ec)) * nhist, dtype=np.uint64)
    ov = C.c_uint64(0)
    lib().or_hist_f32(xyz32, ec, len(ec), np.float32(bin_), int(nhist), h, C.byref(ov),
                      _threads(threads))
    return h.reshape(-1, nhist), int(ov.value)


def debye_pairsum(xyz64, el_count, q, threads=None):
    xyz64, ec, q = _c64(xyz64), _cu32(el_count), _c64(q)
    S = np.zeros(ntypes(len(ec)) * q.size)
    lib().or_debye_pairsum(xyz64, ec, len(ec), q, q.size, S, _threads(threads))
    return S.reshape(-1, q.size)


def partial_debye_pairsum(xyz64, el_count, blk_count, nblocks, q, threads=None):
    xyz64, ec, bc, q = _c64(xyz64), _cu32(el_count), _cu32(blk_count), _c64(q)
    nparts = nblocks * (nblocks + 1) // 2
    S = np.zeros(ntypes(len(ec)) * nparts * q.size)
    lib().or_partial_debye_pairsum(xyz64, ec, len(ec), bc, nblocks, q, q.size, S, _threads(threads))
    return S.reshape(ntypes(len(ec)), nparts, q.size)


def hist_pairsum(hist, bin_, q, threads=None):
    hist = np.ascontiguousarray(hist, dtype=np.uint64)
    q = _c64(q)
    nt, nh = hist.shape
    S = np.zeros(nt * q.size)
    lib().or_hist_pairsum(hist, nt, nh, float(bin_), q, q.size, S, _threads(threads))
    return S.reshape(nt, q.size)


def intensity_1d(S, el_count, q, ff=None, sl=None, polar=False, lam=0.0):
    S, ec, q = _c64(S), _cu32(el_count), _c64(q)
    ffa = None if ff is None else _c64(ff)
    sla = None if sl is None else _c64(sl)
    I = np.zeros(q.size)
    lib().or_intensity_1d(S, len(ec), ec, _ptr(ffa), _ptr(sla), q, q.size, int(polar), float(lam), I)
    return I


def partial_intensity_1d(S, el_count, blk_count, nblocks, q, ff=None, sl=None, polar=False,
                         lam=0.0):
    S, ec, bc, q = _c64(S), _cu32(el_count), _cu32(blk_count), _c64(q)
    ffa = None if ff is None else _c64(ff)
    sla = None if sl is None else _c64(sl)
    nparts = nblocks * (nblocks + 1) // 2
    I = np.zeros((1 + nparts) * q.size)
    lib().or_partial_intensity_1d(S, len(ec), bc, nblocks, _ptr(ffa), _ptr(sla), q, q.size,
                                  int(polar), float(lam), int(ec.sum()), I)
    return I.reshape(1 + nparts, q.size)


def polar_1d(I, q, lam):
    I, q = _c64(I).copy(), _c64(q)
    lib().or_polar_1d(I, q, q.size, float(lam))
    return I


def pdf(hist, el_count, bin_, pdf_type, rho0, sl, cpu_radius_offset=True):
    hist = np.ascontiguousarray(hist, dtype=np.uint64)
    ec, sl = _cu32(el_count), _c64(sl)
    nt, nh = hist.shape
    out = np.zeros((1 + nt) * nh)
    lib().or_pdf(hist, len(ec), ec, nh, float(bin_), int(pdf_type), float(rho0), sl,
                 int(bool(cpu_radius_offset)), out)
    return out.reshape(1 + nt, nh)


def rot_matrix(euler, convention):
    e = _c64(euler)
    R = np.zeros(9)
    lib().or_rot_matrix(e, int(convention), R)
    return R.reshape(3, 3)


def amp2d(xyz64, el_count, qvec, rot, ff, threads=None):
    xyz64, ec = _c64(xyz64), _cu32(el_count)
    qvec = _c64(qvec).reshape(-1, 3)
    rot = _c64(rot).reshape(-1, 9)
    ff = _c64(ff).reshape(len(ec), qvec.shape[0])
    I = np.zeros(qvec.shape[0])
    lib().or_amp2d(xyz64, ec, len(ec), qvec, qvec.shape[0], rot, rot.shape[0], ff, I,
                   _threads(threads))
    return I


def calc_int2d(xyz64, el_count, q, nfi, euler_min, euler_max, euler_n, convention, lam,
               polar=False, ff=None, sl=None, threads=None):
    xyz64, ec, q = _c64(xyz64), _cu32(el_count), _c64(q)
    ffa = None if ff is None else _c64(ff)
    sla = None if sl is None else _c64(sl)
    I2D = np.zeros(q.size * nfi)
    I1D = np.zeros(q.size)
    lib().or_calc_int2d(xyz64, ec, len(ec), q, q.size, int(nfi), _c64(euler_min), _c64(euler_max),
                        _cu32(euler_n), int(convention), float(lam), int(polar), _ptr(ffa),
                        _ptr(sla), I2D, I1D, _threads(threads))
    return I2D.reshape(q.size, nfi), I1D


def philox4x64(ctr, key):
    out = np.zeros(4, np.uint64)
    lib().or_philox4x64(np.ascontiguousarray(ctr, np.uint64), np.ascontiguousarray(key, np.uint64),
                        out)
    return out


class _OrBlock(C.Structure):
    _fields_ = [("natom", C.c_uint), ("nsymm", C.c_uint), ("ncopy", C.c_uint),
                ("ncell", C.c_uint * 3), ("el", C.c_void_p), ("r_atom", C.c_void_p),
                ("occ", C.c_void_p), ("dev", C.c_void_p), ("p_symm", C.c_void_p),
                ("r_symm", C.c_void_p), ("cell", C.c_void_p), ("r_copy", C.c_void_p),
                ("cf", C.c_void_p), ("r_mean", C.c_double * 3), ("rcut", C.c_double),
                ("rcut_copies", C.c_double), ("dev_mol", C.c_double), ("mol_rotation", C.c_int)]


def calc_atoms(blocks, seed):
    """block::calcAtoms over `blocks` (objects with the or_block fields as
    numpy arrays / numbers).  Returns (el, block, xyz[n, 3]) in placement
    order; a stable sort by element gives the reference's ra[el] lists."""
    keep, cs = [], (_OrBlock * len(blocks))()
    cap = 0
    for b, B in zip(blocks, cs):
        arrs = dict(el=_cu32(b.el), r_atom=_c64(b.r_atom), occ=_c64(b.occ), dev=_c64(b.dev),
                    p_symm=_c64(b.p_symm), r_symm=_c64(b.r_symm), cell=_c64(b.cell),
                    r_copy=_c64(b.r_copy), cf=_c64(b.cf))
        keep.append(arrs)
        B.natom, B.nsymm, B.ncopy = int(b.natom), int(b.nsymm), int(b.ncopy)
        B.ncell = (C.c_uint * 3)(*[int(v) for v in b.ncell])
        for k, a in arrs.items():
            setattr(B, k, a.ctypes.data_as(C.c_void_p).value)
        B.r_mean = (C.c_double * 3)(*[float(v) for v in b.r_mean])
        B.rcut, B.rcut_copies, B.dev_mol = float(b.rcut), float(b.rcut_copies), float(b.dev_mol)
        B.mol_rotation = int(b.mol_rotation)
        cap += int(b.natom) * int(b.nsymm) * int(b.ncopy) * int(np.prod(b.ncell))
    el = np.zeros(max(cap, 1), np.uint32)
    blk = np.zeros(max(cap, 1), np.uint32)
    xyz = np.zeros(3 * max(cap, 1))
    n = lib().or_calc_atoms(C.addressof(cs), len(blocks), int(seed) & (2**64 - 1), cap, el, blk, xyz)
    if n < 0:
        raise RuntimeError("or_calc_atoms: more atoms than candidate sites")
    return el[:n], blk[:n], xyz[:3 * n].reshape(n, 3)


def cutoff_pairsum(xyz64, el, n_el, rmax2, q):
    """(S[n_el*n_el, nq], npairs[n_el*n_el]) of or_cutoff_pairsum."""
    xyz64, q = _c64(xyz64), _c64(q)
    S = np.zeros(n_el * n_el * q.size)
    npairs = np.zeros(n_el * n_el, np.uint64)
    lib().or_cutoff_pairsum(xyz64, xyz64.shape[-1], _cu32(el), n_el, _c64(rmax2), q, q.size, S,
                            npairs)
    return S.reshape(n_el * n_el, q.size), npairs


def pairlist_pairsum(r, types, ntypes, q):
    r, q = _c64(r), _c64(q)
    S = np.zeros(ntypes * q.size)
    lib().or_pairlist_pairsum(r, _cu32(types), r.size, ntypes, q, q.size, S)
    return S.reshape(ntypes, q.size)
