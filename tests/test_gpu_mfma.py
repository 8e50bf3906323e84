"""H3 on the int8 matrix cores (XS_PAIRSUM_MFMA): pair sums within 1e-9 of
the FP64 oracle relative to the largest |S| of each histogram, intensities
within 1e-5 pointwise (measured ~1e-9); count-slice counts 1 to 5 (counts up
to 2^34); ragged sizes, q = 0 and a single column."""
import numpy as np
import pytest

import xs_oracle as O
from conftest import rel_err
from xansons_hip import structures as st

pytestmark = pytest.mark.gpu

TOL_I = 1e-5
TOL_S = 1e-9


def mfma_pairsum(ctx, hist, bin_, q):
    return ctx.hist_pairsum(hist, bin_, q, mfma=True)


def check_s(Sg, So, label):
    err = np.abs(Sg - So).max(axis=1) / np.maximum(np.abs(So).max(axis=1), 1e-300)
    print(f"{label}: max |S - S64| / max|S64| = {err.max():.3e}", flush=True)
    assert err.max() <= TOL_S


def test_mfma_fcc_sphere_iq(ctx):
    ens = st.fcc_sphere(radius=14.0)
    x64 = ens.xyz64()
    nh = O.hist_size(x64, 0.004)
    h, _ = O.hist_f64(x64, ens.el_count, 0.004, nh)
    q = st.q_mesh(0.2, 25.0, 700)
    Sg = mfma_pairsum(ctx, h, 0.004, q)
    So = O.hist_pairsum(h, 0.004, q)
    check_s(Sg, So, "fcc")
    pw, l2 = rel_err(O.intensity_1d(Sg, ens.el_count, q, sl=[5.0]),
                     O.intensity_1d(So, ens.el_count, q, sl=[5.0]))
    print(f"fcc I(q): max rel {pw:.3e} L2 {l2:.3e}", flush=True)
    assert pw <= TOL_I and l2 <= TOL_I


def test_mfma_many_columns_one_grid(ctx):
    """40 two-element particles on one 6,000-bin grid: 120 columns."""
    import xansons_hip as xh
    parts = st.c4_particles(40, seed=5, nmin=300, nmax=3000)
    nh = 6000
    b = xh.Batch(ctx, [(p.xyz64(), p.el_count) for p in parts], 0.01, nhist=[nh] * len(parts))
    hist, _, _ = b.run()
    assert not b.overflow().any()
    b.close()
    H = hist.reshape(-1, nh)
    assert H.shape[0] == 120
    q = np.concatenate([[0.0], st.q_mesh(0.3, 20.0, 1000)])
    Sg = mfma_pairsum(ctx, H, 0.01, q)
    So = O.hist_pairsum(H, 0.01, q)
    assert np.all(Sg[:, 0] == 0.0)
    check_s(Sg, So, "120 columns")


@pytest.mark.parametrize("bits", [6, 13, 20, 27, 34])
def test_mfma_count_slices(ctx, bits):
    rng = np.random.Generator(np.random.PCG64(bits))
    nt, nh = 37, 1000 + bits
    H = rng.integers(0, 1 << bits, size=(nt, nh), dtype=np.uint64)
    H[rng.random((nt, nh)) < 0.3] = 0
    q = st.q_mesh(0.1, 30.0, 129)
    Sg = mfma_pairsum(ctx, H, 0.02, q)
    So = O.hist_pairsum(H, 0.02, q)
    check_s(Sg, So, f"counts < 2^{bits}")


def test_mfma_single_column_ragged(ctx):
    ens = st.random_cluster([150, 90], box=9.0, min_dist=0.6, seed=4)
    x64 = ens.xyz64()
    nh = O.hist_size(x64, 0.013)
    h, _ = O.hist_f64(x64, ens.el_count, 0.013, nh)
    q = st.q_mesh(0.5, 12.0, 33)
    Sg = mfma_pairsum(ctx, h[1:2], 0.013, q)
    So = O.hist_pairsum(h[1:2], 0.013, q)
    check_s(Sg, So, "single column")


def test_mfma_rejects_counts_beyond_2_35(ctx):
    import xansons_hip as xh
    H = np.zeros((2, 64), dtype=np.uint64)
    H[1, 3] = np.uint64(1) << np.uint64(35)
    with pytest.raises(xh.XSError):
        mfma_pairsum(ctx, H, 0.01, st.q_mesh(0.5, 5.0, 8))


def test_mfma_full_precision_form_workload(ctx):
    """The precision-form workload at its full size: 256 two-element
    particles, 768 histograms on one 16,384-bin grid (0.004 A), 8,192 q over
    0.2-25 A^-1.  Every column against the FP64 vector kernel (itself pinned
    to the oracle), and 24 columns against an FP64 numpy product."""
    import torch
    import xansons_hip as xh
    nh, bin_, nq = 16384, 0.004, 8192
    parts = st.c4_particles()
    b = xh.Batch(ctx, [(p.xyz64(), p.el_count) for p in parts], bin_, nhist=[nh] * len(parts))
    hist, _, _ = b.run()
    assert not b.overflow().any()
    b.close()
    H = hist.reshape(-1, nh)
    assert H.shape == (768, nh)
    q = st.q_mesh(0.2, 25.0, nq)
    dev = torch.device("cuda", 0)
    hd = torch.from_numpy(H.view(np.int64)).to(dev)
    qd = torch.from_numpy(q).to(dev)
    Sv = torch.empty(768 * nq, dtype=torch.float64, device=dev)
    Sm = torch.empty_like(Sv)
    torch.cuda.synchronize()
    print("c4 full: 768 x 16384 x 8192", flush=True)
    ctx.hist_pairsum_dev(hd.data_ptr(), 768, nh, bin_, qd.data_ptr(), nq, Sv.data_ptr())
    ctx.hist_pairsum_dev(hd.data_ptr(), 768, nh, bin_, qd.data_ptr(), nq, Sm.data_ptr(), mfma=True)
    ctx.synchronize()
    Sv = Sv.view(768, nq).cpu().numpy()
    Sm = Sm.view(768, nq).cpu().numpy()
    check_s(Sm, Sv, "c4 all columns, MFMA vs FP64 vector")
    cols = np.arange(0, 768, 32)
    r = (np.arange(nh) + 0.5) * bin_
    qr = np.outer(q, r)
    A = np.sin(qr) / (qr + 1e-8)
    del qr
    Se = (A @ H[cols].astype(np.float64).T).T
    del A
    check_s(Sm[cols], Se, "c4 24 columns, MFMA vs FP64 numpy")
    check_s(Sv[cols], Se, "c4 24 columns, FP64 vector vs FP64 numpy")
