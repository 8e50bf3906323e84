"""Python binding of libxs_hip.so, the gfx950 backend of the XaNSoNS calculators.

The functions here call the C ABI declared in include/xs_hip.h through
ctypes.  There is no CPU fallback: if the shared library is missing, or no
GPU is visible, the calls raise.

Host-array functions (numpy in, numpy out) mirror the reference's
calculators; ``*_dev`` variants take device pointers (ints, e.g. from
``torch.Tensor.data_ptr()``) and enqueue on the context stream.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import weakref
from typing import Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libxs_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "xs_hip.h")

XS_OK = 0
XS_DEVICE_PTRS = 0x1
XS_F64_BINNING = 0x2
XS_PDF_CPU_RADIUS = 0x4
XS_BATCH_GRAPH = 0x8
XS_PAIRSUM_MFMA = 0x10
XS_F32_DOUBLE_QUOTIENT = 0x20
XS_PAIRSUM_F32 = 0x40
PDF_TYPES = {"RDF": 0, "PDF": 1, "rPDF": 2}

_lib = None


class XSError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"xs error {code}: {msg}")
        self.code = code


def build(quiet: bool = True) -> None:
    """Compile libxs_hip.so in-tree with hipcc for gfx950."""
    cmd = ["make", "-C", PKG_DIR, "-j8"]
    subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL if quiet else None)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, sz, u, d, i = C.c_void_p, C.c_size_t, C.c_uint, C.c_double, C.c_int
        L.xs_last_error.restype = C.c_char_p
        L.xs_init.argtypes = [i, C.POINTER(vp)]
        L.xs_free.argtypes = [vp]
        L.xs_set_stream.argtypes = [vp, vp]
        L.xs_synchronize.argtypes = [vp]
        L.xs_device_count.argtypes = [C.POINTER(i)]
        L.xs_profile.argtypes = [vp, i]
        L.xs_profile_read.argtypes = [vp, C.c_char_p, C.POINTER(d), C.POINTER(C.c_ulonglong)]
        L.xs_hist.argtypes = [vp, vp, sz, vp, u, d, u, vp, vp, u]
        L.xs_hist_pairsum.argtypes = [vp, vp, u, u, d, vp, u, vp, u]
        L.xs_pdf.argtypes = [vp, vp, u, vp, u, d, i, d, vp, vp, u]
        L.xs_debye_pairsum.argtypes = [vp, vp, sz, vp, u, vp, u, vp, u]
        L.xs_partial_debye_pairsum.argtypes = [vp, vp, sz, vp, u, vp, u, vp, u, vp, u]
        L.xs_debye_pairsum_noalloc.argtypes = [vp, vp, sz, vp, u, vp, u, vp, u]
        L.xs_amp2d.argtypes = [vp, vp, sz, vp, u, vp, sz, vp, u, vp, vp, u]
        L.xs_batch_plan.argtypes = [vp, u, vp, vp, vp, d, vp, u, C.POINTER(vp)]
        L.xs_batch_info.argtypes = [vp, vp, vp, vp, vp, vp]
        L.xs_batch_run.argtypes = [vp, vp, vp, u, vp, i, vp, vp, vp, u]
        L.xs_batch_overflow.argtypes = [vp, vp]
        L.xs_batch_free.argtypes = [vp]
        L.xs_amp2d_grid.argtypes = [vp, vp, sz, vp, u, vp, vp, vp, u, u, vp, u, vp, vp, u]
        L.xs_cutoff_pairsum.argtypes = [vp, vp, sz, vp, u, vp, vp, u, vp, vp, u]
        L.xs_pairlist_pairsum.argtypes = [vp, vp, vp, sz, u, vp, u, vp, u]
        L.xs_atoms_build.argtypes = [vp, vp, u, u, C.c_uint64, C.POINTER(vp)]
        L.xs_atoms_info.argtypes = [vp, vp, vp, vp]
        L.xs_atoms_copy.argtypes = [vp, vp, u]
        L.xs_atoms_free.argtypes = [vp]
        _lib = L
    return _lib


def header_symbols(path: str = HEADER_PATH) -> list:
    """Function names declared in include/xs_hip.h."""
    import re
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(xs_\w+)\s*\(", text, re.M)))


def _check(code: int) -> None:
    if code != XS_OK:
        raise XSError(code, lib().xs_last_error().decode())


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


def ntypes(n_el: int) -> int:
    return n_el * (n_el + 1) // 2


class Context:
    """One xs_ctx (device, stream, workspace)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        self._batches = weakref.WeakSet()
        self._atoms = weakref.WeakSet()
        _check(lib().xs_init(int(device), C.byref(self.h)))

    def close(self):
        """Free the context; batch plans and atom sets made on it are freed first."""
        for b in list(self._batches) + list(self._atoms):
            b.close()
        if self.h:
            lib().xs_free(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self):
        _check(lib().xs_synchronize(self.h))

    def set_stream(self, stream_ptr: Optional[int]):
        _check(lib().xs_set_stream(self.h, C.c_void_p(stream_ptr) if stream_ptr else None))

    def profile(self, enable: bool = True):
        _check(lib().xs_profile(self.h, int(enable)))

    def profile_read(self, kernel: str):
        ms, n = C.c_double(0), C.c_ulonglong(0)
        _check(lib().xs_profile_read(self.h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---------------------------------------------------------------- H1
    def hist(self, xyz, el_count, bin_: float, nhist: int, f64: bool = False,
             double_quotient: bool = False):
        """Histogram from host coordinates (3, N).  f64=True bins in double as
        the reference CPU path (xyz promoted to float64), else in float;
        double_quotient keeps the FP32 binning's quotient as a rounded double
        product (same counts)."""
        ec = _u32(el_count)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64 if f64 else np.float32)
        n = int(xyz.shape[-1]) if xyz.ndim > 1 else xyz.size // 3
        h = np.zeros(ntypes(len(ec)) * int(nhist), dtype=np.uint64)
        ov = np.zeros(1, dtype=np.uint64)
        flags = (XS_F64_BINNING if f64 else 0) | (XS_F32_DOUBLE_QUOTIENT if double_quotient else 0)
        _check(lib().xs_hist(self.h, _p(xyz), n, _p(ec), len(ec), float(bin_), int(nhist), _p(h),
                             _p(ov), flags))
        return h.reshape(-1, int(nhist)), int(ov[0])

    def hist_dev(self, xyz_ptr: int, n: int, el_count, bin_: float, nhist: int, hist_ptr: int,
                 ovf_ptr: Optional[int] = None, f64: bool = False, double_quotient: bool = False):
        ec = _u32(el_count)
        flags = XS_DEVICE_PTRS | (XS_F64_BINNING if f64 else 0) | \
            (XS_F32_DOUBLE_QUOTIENT if double_quotient else 0)
        _check(lib().xs_hist(self.h, C.c_void_p(xyz_ptr), int(n), _p(ec), len(ec), float(bin_),
                             int(nhist), C.c_void_p(hist_ptr),
                             C.c_void_p(ovf_ptr) if ovf_ptr else None, flags))

    # ---------------------------------------------------------------- H3
    def hist_pairsum(self, hist, bin_: float, q, mfma: bool = False, f32: bool = False):
        """mfma=True: the int8 matrix-core product (many histograms); f32=True:
        the plain FP32 form (comparison only)."""
        hist = np.ascontiguousarray(hist, dtype=np.uint64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        nt, nh = hist.shape
        S = np.zeros(nt * q.size)
        _check(lib().xs_hist_pairsum(self.h, _p(hist), nt, nh, float(bin_), _p(q), q.size, _p(S),
                                     (XS_PAIRSUM_MFMA if mfma else 0) | (XS_PAIRSUM_F32 if f32 else 0)))
        return S.reshape(nt, q.size)

    def hist_pairsum_dev(self, hist_ptr, ntypes_, nhist, bin_, q_ptr, nq, S_ptr, mfma=False,
                         f32=False):
        """f32=True: the plain FP32 form (comparison only)."""
        _check(lib().xs_hist_pairsum(self.h, C.c_void_p(hist_ptr), int(ntypes_), int(nhist),
                                     float(bin_), C.c_void_p(q_ptr), int(nq), C.c_void_p(S_ptr),
                                     XS_DEVICE_PTRS | (XS_PAIRSUM_MFMA if mfma else 0) |
                                     (XS_PAIRSUM_F32 if f32 else 0)))

    # ---------------------------------------------------------------- H4
    def pdf(self, hist, el_count, bin_, pdf_type, rho0, sl, cpu_radius=False):
        hist = np.ascontiguousarray(hist, dtype=np.uint64)
        ec, sl = _u32(el_count), np.ascontiguousarray(sl, dtype=np.float64)
        nt, nh = hist.shape
        out = np.zeros((1 + nt) * nh)
        t = PDF_TYPES.get(pdf_type, pdf_type)
        flags = XS_PDF_CPU_RADIUS if cpu_radius else 0
        _check(lib().xs_pdf(self.h, _p(hist), len(ec), _p(ec), nh, float(bin_), int(t), float(rho0),
                            _p(sl), _p(out), flags))
        return out.reshape(1 + nt, nh)

    def pdf_dev(self, hist_ptr, el_count, nhist, bin_, pdf_type, rho0, sl, out_ptr,
                cpu_radius=False):
        ec, sl = _u32(el_count), np.ascontiguousarray(sl, dtype=np.float64)
        t = PDF_TYPES.get(pdf_type, pdf_type)
        flags = XS_DEVICE_PTRS | (XS_PDF_CPU_RADIUS if cpu_radius else 0)
        _check(lib().xs_pdf(self.h, C.c_void_p(hist_ptr), len(ec), _p(ec), int(nhist), float(bin_),
                            int(t), float(rho0), _p(sl), C.c_void_p(out_ptr), flags))

    # ---------------------------------------------------------------- H5/H6
    def debye_pairsum(self, xyz, el_count, q):
        ec = _u32(el_count)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        S = np.zeros(ntypes(len(ec)) * q.size)
        _check(lib().xs_debye_pairsum(self.h, _p(xyz), int(ec.sum()), _p(ec), len(ec), _p(q),
                                      q.size, _p(S), 0))
        return S.reshape(-1, q.size)

    def debye_pairsum_dev(self, xyz_ptr, n, el_count, q_ptr, nq, S_ptr):
        ec = _u32(el_count)
        _check(lib().xs_debye_pairsum(self.h, C.c_void_p(xyz_ptr), int(n), _p(ec), len(ec),
                                      C.c_void_p(q_ptr), int(nq), C.c_void_p(S_ptr), XS_DEVICE_PTRS))

    def debye_pairsum_noalloc_dev(self, xyz_ptr, n, el_count, q_ptr, nq, S_ptr):
        """xs_debye_pairsum_noalloc: device pointers, no allocation, no wait."""
        ec = _u32(el_count)
        _check(lib().xs_debye_pairsum_noalloc(self.h, C.c_void_p(xyz_ptr), int(n), _p(ec), len(ec),
                                              C.c_void_p(q_ptr), int(nq), C.c_void_p(S_ptr),
                                              XS_DEVICE_PTRS))

    def partial_debye_pairsum(self, xyz, el_count, blk_count, nblocks, q):
        ec, bc = _u32(el_count), _u32(blk_count)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        nparts = nblocks * (nblocks + 1) // 2
        S = np.zeros(ntypes(len(ec)) * nparts * q.size)
        _check(lib().xs_partial_debye_pairsum(self.h, _p(xyz), int(ec.sum()), _p(ec), len(ec),
                                              _p(bc), int(nblocks), _p(q), q.size, _p(S), 0))
        return S.reshape(ntypes(len(ec)), nparts, q.size)

    # ---------------------------------------------------------------- H7
    def cutoff_pairsum(self, xyz, el, n_el: int, rmax2, q):
        """RearrangementInt's cut-off sums: (S[n_el*n_el, nq], npairs[n_el*n_el])."""
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        ea, r2, qa = _u32(el), np.ascontiguousarray(rmax2, np.float64), \
            np.ascontiguousarray(q, np.float64)
        S = np.zeros(n_el * n_el * qa.size)
        npairs = np.zeros(n_el * n_el, dtype=np.uint64)
        _check(lib().xs_cutoff_pairsum(self.h, _p(x), x.shape[-1], _p(ea), int(n_el), _p(r2),
                                       _p(qa), qa.size, _p(S), _p(npairs), 0))
        return S.reshape(n_el * n_el, qa.size), npairs

    def pairlist_pairsum(self, r, types, ntypes: int, q):
        """Sine sums over a distance list: S[ntypes, nq]."""
        ra, ta, qa = np.ascontiguousarray(r, np.float64), _u32(types), \
            np.ascontiguousarray(q, np.float64)
        S = np.zeros(ntypes * qa.size)
        _check(lib().xs_pairlist_pairsum(self.h, _p(ra), _p(ta), ra.size, int(ntypes), _p(qa),
                                         qa.size, _p(S), 0))
        return S.reshape(ntypes, qa.size)

    def amp2d(self, xyz, el_count, qvec, rot, ff):
        ec = _u32(el_count)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        qvec = np.ascontiguousarray(qvec, dtype=np.float64).reshape(-1, 3)
        rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(-1, 9)
        ff = np.ascontiguousarray(ff, dtype=np.float64).reshape(len(ec), qvec.shape[0])
        I = np.zeros(qvec.shape[0])
        _check(lib().xs_amp2d(self.h, _p(xyz), int(ec.sum()), _p(ec), len(ec), _p(qvec),
                              qvec.shape[0], _p(rot), rot.shape[0], _p(ff), _p(I), 0))
        return I

    def amp2d_grid(self, xyz, el_count, q0, du, dv, nu: int, nv: int, rot, ff):
        """xs_amp2d_grid: q(iu, iv) = q0 + iu du + iv dv, point k = iv nu + iu."""
        ec = _u32(el_count)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(-1, 9)
        ff = np.ascontiguousarray(ff, dtype=np.float64).reshape(len(ec), nu * nv)
        v3 = [np.ascontiguousarray(v, dtype=np.float64).reshape(3) for v in (q0, du, dv)]
        I = np.zeros(nu * nv)
        _check(lib().xs_amp2d_grid(self.h, _p(xyz), int(ec.sum()), _p(ec), len(ec), _p(v3[0]),
                                   _p(v3[1]), _p(v3[2]), int(nu), int(nv), _p(rot), rot.shape[0],
                                   _p(ff), _p(I), 0))
        return I

    def amp2d_grid_dev(self, xyz_ptr, n, el_count, q0, du, dv, nu, nv, rot_ptr, nrot, ff_ptr,
                       I_ptr):
        ec = _u32(el_count)
        v3 = [np.ascontiguousarray(v, dtype=np.float64).reshape(3) for v in (q0, du, dv)]
        _check(lib().xs_amp2d_grid(self.h, C.c_void_p(xyz_ptr), int(n), _p(ec), len(ec),
                                   _p(v3[0]), _p(v3[1]), _p(v3[2]), int(nu), int(nv),
                                   C.c_void_p(rot_ptr), int(nrot), C.c_void_p(ff_ptr),
                                   C.c_void_p(I_ptr), XS_DEVICE_PTRS))

    def amp2d_dev(self, xyz_ptr, n, el_count, qvec_ptr, nqp, rot_ptr, nrot, ff_ptr, I_ptr):
        ec = _u32(el_count)
        _check(lib().xs_amp2d(self.h, C.c_void_p(xyz_ptr), int(n), _p(ec), len(ec),
                              C.c_void_p(qvec_ptr), int(nqp), C.c_void_p(rot_ptr), int(nrot),
                              C.c_void_p(ff_ptr), C.c_void_p(I_ptr), XS_DEVICE_PTRS))


class Batch:
    """A batch plan (xs_batch_plan) over an ensemble of structures.

    ``structures`` is a list of (xyz (3, N) array, el_count) pairs, each
    element-major.  ``f64`` bins in double as the reference CPU path."""

    def __init__(self, ctx: Context, structures, bin_: float, nhist=None, f64: bool = True,
                 double_quotient: bool = False, device: bool = False):
        """device=True: ``structures`` holds (device pointer of the (3, N)
        float64 SoA, el_count) pairs and the plan is tiled on the GPU."""
        self.ctx = ctx
        self.bin = float(bin_)
        n_el = _u32([len(ec) for _, ec in structures])
        ecs = _u32(np.concatenate([np.asarray(ec, dtype=np.uint32) for _, ec in structures]))
        if device:
            self._xyz = None
            ptrs = (C.c_void_p * len(structures))(*[int(x) for x, _ in structures])
        else:
            self._xyz = [np.ascontiguousarray(x, dtype=np.float64) for x, _ in structures]
            ptrs = (C.c_void_p * len(structures))(*[x.ctypes.data for x in self._xyz])
        nh = None if nhist is None else _u32(nhist)
        self.h = C.c_void_p()
        flags = (XS_F64_BINNING if f64 else 0) | (XS_F32_DOUBLE_QUOTIENT if double_quotient else 0) | \
            (XS_DEVICE_PTRS if device else 0)
        _check(lib().xs_batch_plan(ctx.h, len(structures), _p(n_el), _p(ecs), ptrs, self.bin,
                                   _p(nh), flags, C.byref(self.h)))
        self._xyz = None
        ctx._batches.add(self)
        self.nstruct = len(structures)
        self.n_el = [int(v) for v in n_el]
        self.el_count = ecs
        ns = self.nstruct
        self.nhist = np.zeros(ns, dtype=np.uint32)
        self.hist_off = np.zeros(ns, dtype=np.uint64)
        self.pdf_off = np.zeros(ns, dtype=np.uint64)
        self.type_first = np.zeros(ns, dtype=np.uint64)
        tot = np.zeros(3, dtype=np.uint64)
        _check(lib().xs_batch_info(self.h, _p(self.nhist), _p(self.hist_off), _p(self.pdf_off),
                                   _p(self.type_first), _p(tot)))
        self.hist_len, self.pdf_len, self.ntypes_total = (int(v) for v in tot)

    def close(self):
        if self.h:
            lib().xs_batch_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, q=None, pdf_type=None, rho0=None, sl=None, cpu_radius=True, graph=False,
            mfma=False):
        """Host arrays in and out: (hist, S or None, pdf or None) as flat arrays.
        mfma=True: the pair sums on the int8 matrix cores."""
        hist = np.zeros(self.hist_len, dtype=np.uint64)
        S = pdf = None
        qa = None
        if q is not None:
            qa = np.ascontiguousarray(q, dtype=np.float64)
            S = np.zeros(self.ntypes_total * qa.size)
        t = -1
        rho = sla = None
        if pdf_type is not None:
            t = PDF_TYPES.get(pdf_type, pdf_type)
            pdf = np.zeros(self.pdf_len)
            sla = np.ascontiguousarray(sl, dtype=np.float64)
            rho = None if rho0 is None else np.ascontiguousarray(rho0, dtype=np.float64)
        flags = (XS_PDF_CPU_RADIUS if cpu_radius else 0) | (XS_BATCH_GRAPH if graph else 0) | \
            (XS_PAIRSUM_MFMA if mfma else 0)
        _check(lib().xs_batch_run(self.h, _p(hist), _p(qa), 0 if qa is None else qa.size, _p(S),
                                  int(t), _p(rho), _p(sla), _p(pdf), flags))
        return hist, S, pdf

    def run_dev(self, hist_ptr, q_ptr=None, nq=0, S_ptr=None, pdf_type=None, rho0=None, sl=None,
                pdf_ptr=None, cpu_radius=True, graph=False, mfma=False):
        t = -1 if pdf_type is None else PDF_TYPES.get(pdf_type, pdf_type)
        sla = None if sl is None else np.ascontiguousarray(sl, dtype=np.float64)
        rho = None if rho0 is None else np.ascontiguousarray(rho0, dtype=np.float64)
        flags = XS_DEVICE_PTRS | (XS_PDF_CPU_RADIUS if cpu_radius else 0) | \
            (XS_BATCH_GRAPH if graph else 0) | (XS_PAIRSUM_MFMA if mfma else 0)
        vp = lambda v: C.c_void_p(v) if v else None  # noqa: E731
        _check(lib().xs_batch_run(self.h, vp(hist_ptr), vp(q_ptr), int(nq), vp(S_ptr), int(t),
                                  _p(rho), _p(sla), vp(pdf_ptr), flags))

    def overflow(self):
        o = np.zeros(self.nstruct, dtype=np.uint64)
        _check(lib().xs_batch_overflow(self.h, _p(o)))
        return o


def device_count() -> int:
    n = C.c_int(0)
    code = lib().xs_device_count(C.byref(n))
    return n.value if code == XS_OK else 0
