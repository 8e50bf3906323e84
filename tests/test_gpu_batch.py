"""Ensemble runs (xs_batch_plan / xs_batch_run): every structure's histograms
equal the oracle's bit for bit (FP64 and FP32 binning), PDFs equal the
oracle's bit for bit, intensities from the pair sums lie within 1e-5 of the
FP64 oracle, and graph replays reproduce the eager run."""
import numpy as np
import pytest

import xs_oracle as O
from conftest import rel_err
from xansons_hip import structures as st

pytestmark = pytest.mark.gpu

TOL = 1e-5
BIN = 0.01


@pytest.fixture(scope="module")
def ensemble():
    # 9 particles of 300-6000 atoms (1-4 elements; both histogram window
    # classes: nhist below and above 8192) plus a two-atom structure
    ens = st.particle_ensemble(9, seed=21, nmin=300, nmax=6000)
    two = st.Ensemble(np.array([[0.0, 1.234], [0.0, 0.0], [0.0, 0.0]], dtype=np.float32), [1, 1],
                      ["A", "B"])
    return ens + [two]


@pytest.fixture(scope="module")
def bctx():
    import xansons_hip as xh
    c = xh.Context(0)
    yield c
    c.close()


def sl_for(ens):
    return np.concatenate([np.linspace(2.0, 9.0, e.n_el) for e in ens])


def rho_for(ens):
    return np.array([O.atomic_density(e.xyz64()) for e in ens])


def split(b, ens, hist, S, pdf, nq):
    out = []
    for s, e in enumerate(ens):
        P = e.n_el * (e.n_el + 1) // 2
        nh = int(b.nhist[s])
        h = hist[int(b.hist_off[s]):int(b.hist_off[s]) + P * nh].reshape(P, nh)
        Ss = None if S is None else S[int(b.type_first[s]) * nq:(int(b.type_first[s]) + P) * nq]
        ps = None if pdf is None else pdf[int(b.pdf_off[s]):int(b.pdf_off[s]) + (1 + P) * nh]
        out.append((h, None if Ss is None else Ss.reshape(P, nq),
                    None if ps is None else ps.reshape(1 + P, nh)))
    return out


def test_batch_f64_matches_oracle(bctx, ensemble):
    import xansons_hip as xh
    q = st.q_mesh(0.5, 20.0, 600)
    sl, rho = sl_for(ensemble), rho_for(ensemble)
    b = xh.Batch(bctx, [(e.xyz64(), e.el_count) for e in ensemble], BIN, f64=True)
    assert max(b.nhist) > 8192 >= min(b.nhist)
    hist, S, pdf = b.run(q=q, pdf_type="rPDF", rho0=rho, sl=sl)
    assert not b.overflow().any()
    so = 0
    for s, (e, (h, Ss, ps)) in enumerate(zip(ensemble, split(b, ensemble, hist, S, pdf, q.size))):
        x64 = e.xyz64()
        nh = O.hist_size(x64, BIN)
        assert int(b.nhist[s]) == nh
        ho, ovf = O.hist_f64(x64, e.el_count, BIN, nh)
        assert ovf == 0
        mism = int(np.count_nonzero(h != ho))
        print(f"struct {s}: N={e.n} n_el={e.n_el} nhist={nh} mismatches={mism}", flush=True)
        assert mism == 0
        sle = sl[so:so + e.n_el]
        so += e.n_el
        po = O.pdf(ho, e.el_count, BIN, xh.PDF_TYPES["rPDF"], rho[s], sle, cpu_radius_offset=True)
        assert np.array_equal(ps, po)
        Io = O.intensity_1d(O.hist_pairsum(ho, BIN, q), e.el_count, q, sl=sle)
        Ig = O.intensity_1d(Ss, e.el_count, q, sl=sle)
        pw, l2 = rel_err(Ig, Io)
        assert pw <= TOL and l2 <= TOL, (s, pw, l2)
    b.close()


@pytest.mark.parametrize("double_quotient", [False, True])
def test_batch_f32_matches_oracle(bctx, ensemble, double_quotient):
    """FP32 binning over the ensemble, with the checked FP32 quotient form and
    with the double product: both equal the oracle."""
    import xansons_hip as xh
    b = xh.Batch(bctx, [(e.xyz, e.el_count) for e in ensemble], BIN, f64=False,
                 double_quotient=double_quotient)
    hist, _, _ = b.run()
    for s, (e, (h, _, _)) in enumerate(zip(ensemble, split(b, ensemble, hist, None, None, 0))):
        ho, _ = O.hist_f32(e.xyz, e.el_count, BIN, int(b.nhist[s]))
        assert int(np.count_nonzero(h != ho)) == 0, s
    b.close()


def device_plan(bctx, ens, **kw):
    """A plan tiled on the GPU from device copies of the coordinates."""
    import torch
    import xansons_hip as xh
    dev = torch.device("cuda", 0)
    xd = [torch.from_numpy(np.ascontiguousarray(e.xyz64())).to(dev) for e in ens]
    torch.cuda.synchronize()
    b = xh.Batch(bctx, [(x.data_ptr(), e.el_count) for x, e in zip(xd, ens)], BIN, device=True, **kw)
    del xd  # the plan keeps no reference to the input coordinates
    return b


@pytest.mark.parametrize("f64", [True, False])
def test_device_plan_equals_host_plan(bctx, ensemble, f64):
    """Tiling on the GPU (radix sorts over all structures) gives the same
    bin counts, PDFs and pair sums as host tiling: the counts do not depend
    on the tiling, and both equal the oracle."""
    import xansons_hip as xh
    q = st.q_mesh(0.5, 20.0, 300)
    sl, rho = sl_for(ensemble), rho_for(ensemble)
    bh = xh.Batch(bctx, [(e.xyz64(), e.el_count) for e in ensemble], BIN, f64=f64)
    bd = device_plan(bctx, ensemble, f64=f64)
    assert list(bd.nhist) == list(bh.nhist)
    assert (bd.hist_off == bh.hist_off).all() and (bd.pdf_off == bh.pdf_off).all()
    hh, Sh, ph = bh.run(q=q, pdf_type="rPDF", rho0=rho, sl=sl)
    hd, Sd, pd = bd.run(q=q, pdf_type="rPDF", rho0=rho, sl=sl)
    print(f"device plan f64={f64}: mismatched bins {int(np.count_nonzero(hh != hd))}", flush=True)
    assert np.array_equal(hh, hd)
    assert np.array_equal(ph, pd)
    assert np.allclose(Sd, Sh, rtol=1e-12, atol=1e-12 * np.abs(Sh).max())
    for s, (e, (h, _, _)) in enumerate(zip(ensemble, split(bd, ensemble, hd, None, None, 0))):
        ho, _ = (O.hist_f64(e.xyz64(), e.el_count, BIN, int(bd.nhist[s])) if f64 else
                 O.hist_f32(e.xyz, e.el_count, BIN, int(bd.nhist[s])))
        assert np.array_equal(h, ho), s
    bh.close()
    bd.close()


def test_device_plan_wide_runs_fall_back_to_kd_tiles(bctx):
    """Structures whose equal-count tiles would be wider than the window
    allows: two clusters 160 A apart along z with unequal counts (one
    column; the run that holds the last atoms of one cluster and the first
    of the other spans the gap), and a hollow shell.  Such a structure is
    re-tiled with the host k-d split.  Counts equal the oracle, beside a
    solid sphere tiled on the GPU."""
    rng = np.random.default_rng(7)
    a = st.fcc_sphere(radius=20.0, seed=3).xyz64()
    b = st.fcc_sphere(radius=16.0, seed=4).xyz64()
    two = np.concatenate([a + np.array([[0.0], [0.0], [-80.0]]), b + np.array([[0.0], [0.0], [80.0]])],
                         axis=1)
    u = rng.normal(size=(3, 30000))
    shell = u / np.linalg.norm(u, axis=0) * 60.0 + rng.normal(0, 0.3, (3, 30000))
    ens = [st.Ensemble(np.ascontiguousarray(two), [two.shape[1]], ["A"]),
           st.Ensemble(np.ascontiguousarray(shell), [30000], ["A"]),
           st.Ensemble(np.ascontiguousarray(a), [a.shape[1]], ["B"]),
           st.Ensemble(np.ascontiguousarray(np.concatenate([a, shell], axis=1)),
                       [a.shape[1], 30000], ["B", "A"])]
    bd = device_plan(bctx, ens, f64=True)
    assert int(bd.nhist[0]) > 8192 and int(bd.nhist[1]) > 8192
    hd, _, _ = bd.run()
    for s, (e, (h, _, _)) in enumerate(zip(ens, split(bd, ens, hd, None, None, 0))):
        ho, _ = O.hist_f64(e.xyz64(), e.el_count, BIN, int(bd.nhist[s]))
        print(f"wide-run case {s}: N={e.n} nhist={int(bd.nhist[s])} "
              f"mismatches={int(np.count_nonzero(h != ho))}", flush=True)
        assert np.array_equal(h, ho), s
    bd.close()


def test_batch_given_nhist_and_overflow(bctx, ensemble):
    import xansons_hip as xh
    sub = ensemble[:4]
    nh = [max(50, O.hist_size(e.xyz64(), BIN) // 3) for e in sub]
    b = xh.Batch(bctx, [(e.xyz64(), e.el_count) for e in sub], BIN, nhist=nh, f64=True)
    assert list(b.nhist) == nh
    hist, _, _ = b.run()
    ov = b.overflow()
    for s, (e, (h, _, _)) in enumerate(zip(sub, split(b, sub, hist, None, None, 0))):
        ho, ovo = O.hist_f64(e.xyz64(), e.el_count, BIN, nh[s])
        assert np.array_equal(h, ho)
        assert int(ov[s]) == ovo > 0
    b.close()


@pytest.mark.parametrize("device", [False, True])
def test_batch_grids_beyond_fixed_point_limit(bctx, ensemble, device):
    """A plan whose largest grid exceeds 2^19 bins runs the FP64 kernel with
    the fract form of the fast path for all its structures (the fixed-point
    form takes grids of up to 2^19 bins): given bin counts of 2^19 + 7 and
    1,000-3,000 over four structures, host and device plans, bit-exact."""
    import torch
    import xansons_hip as xh
    sub = ensemble[:4]
    nh = [(1 << 19) + 7, 1000, 2000, 3000]
    if device:
        dev = torch.device("cuda", 0)
        keep = [torch.from_numpy(np.ascontiguousarray(e.xyz64())).to(dev) for e in sub]
        b = xh.Batch(bctx, [(t.data_ptr(), e.el_count) for t, e in zip(keep, sub)], BIN, nhist=nh,
                     f64=True, device=True)
    else:
        b = xh.Batch(bctx, [(e.xyz64(), e.el_count) for e in sub], BIN, nhist=nh, f64=True)
    hist, _, _ = b.run()
    ov = b.overflow()
    for s, (e, (h, _, _)) in enumerate(zip(sub, split(b, sub, hist, None, None, 0))):
        ho, ovo = O.hist_f64(e.xyz64(), e.el_count, BIN, nh[s])
        print(f"struct {s}: nhist={nh[s]} mismatches={int(np.count_nonzero(h != ho))}", flush=True)
        assert np.array_equal(h, ho), s
        assert int(ov[s]) == ovo
    b.close()


def test_batch_graph_replays_match_eager(bctx, ensemble):
    import torch
    import xansons_hip as xh
    dev = torch.device("cuda", 0)
    q = st.q_mesh(0.5, 20.0, 512)
    sl, rho = sl_for(ensemble), rho_for(ensemble)
    b = xh.Batch(bctx, [(e.xyz64(), e.el_count) for e in ensemble], BIN, f64=True)
    h_ref, S_ref, p_ref = b.run(q=q, pdf_type="PDF", rho0=rho, sl=sl)
    qd = torch.from_numpy(q).to(dev)
    hist = torch.empty(b.hist_len, dtype=torch.int64, device=dev)
    S = torch.empty(b.ntypes_total * q.size, dtype=torch.float64, device=dev)
    pdf = torch.empty(b.pdf_len, dtype=torch.float64, device=dev)
    h2 = torch.empty_like(hist)
    for rep in range(4):
        hist.fill_(-1)
        S.fill_(float("nan"))
        pdf.fill_(-1.0)
        torch.cuda.synchronize()
        b.run_dev(hist.data_ptr(), qd.data_ptr(), q.size, S.data_ptr(), "PDF", rho, sl, pdf.data_ptr(),
                  graph=True)
        bctx.synchronize()
        hh = hist.cpu().numpy().view(np.uint64)
        assert np.array_equal(hh, h_ref), rep
        assert np.array_equal(pdf.cpu().numpy(), p_ref, equal_nan=True), rep  # rho0 = 0: 0/0
        Sg = S.cpu().numpy()
        assert np.allclose(Sg, S_ref, rtol=1e-12, atol=1e-9 * np.abs(S_ref).max()), rep
        # eager work of the program between replays: another (non-graph) run
        b.run_dev(h2.data_ptr())
        bctx.synchronize()
        assert np.array_equal(h2.cpu().numpy().view(np.uint64), h_ref)
        print(f"replay {rep}: ok", flush=True)
    b.close()


def test_batch_full_ensemble_workload(bctx):
    """The ensemble workload at its full size: 1,000 particles of 500-200,000
    atoms (1-4 elements), bin 0.01 A, planned on the GPU.  The batched
    histograms equal the
    per-structure xs_hist calls bit for bit for every particle (both FP64
    binning; the per-structure path is pinned to the oracle by the other
    tests), the 40 smallest particles and the smallest of over 20,000
    atoms equal the oracle directly, and pair totals are exact for all.
    Pair sums on 2,048 q: the FP64 kernel's I(q) within 1e-5 of the oracle
    on that sample, the matrix-core form's within 1e-5 of the FP64 kernel's
    for every particle."""
    import torch
    import xansons_hip as xh
    parts = st.particle_ensemble(1000)
    b = device_plan(bctx, parts, f64=True)
    dev = torch.device("cuda", 0)
    hist = torch.empty(b.hist_len, dtype=torch.int64, device=dev)
    hist2 = torch.empty_like(hist)
    torch.cuda.synchronize()
    print("full ensemble: batched run", flush=True)
    b.run_dev(hist.data_ptr())
    bctx.synchronize()
    assert not b.overflow().any()
    print("full ensemble: per-structure runs", flush=True)
    for s, p in enumerate(parts):
        P = p.n_el * (p.n_el + 1) // 2
        xd = torch.from_numpy(p.xyz64()).to(dev)
        bctx.hist_dev(xd.data_ptr(), p.n, p.el_count, BIN, int(b.nhist[s]),
                      hist2.data_ptr() + 8 * int(b.hist_off[s]), None, f64=True)
        if s % 100 == 99:
            bctx.synchronize()
    bctx.synchronize()
    assert torch.equal(hist, hist2)
    H = hist.cpu().numpy().view(np.uint64)
    sizes = np.array([p.n for p in parts])
    big = np.flatnonzero(sizes > 20000)
    sample = list(np.argsort(sizes, kind="stable")[:40]) + [int(big[np.argmin(sizes[big])])]
    for s, p in enumerate(parts):
        P = p.n_el * (p.n_el + 1) // 2
        nh = int(b.nhist[s])
        h = H[int(b.hist_off[s]):int(b.hist_off[s]) + P * nh].reshape(P, nh)
        ec = list(p.el_count)
        want = [ec[a] * (ec[a] - 1) // 2 if a == c else ec[a] * ec[c]
                for a in range(p.n_el) for c in range(a, p.n_el)]
        assert [int(v) for v in h.sum(axis=1)] == want
        if s in sample:
            o, _ = O.hist_f64(p.xyz64(), ec, BIN, nh)
            assert np.array_equal(h, o), s
    # pair sums on 2,048 q: the FP64 kernel against the oracle on the
    # sample, the matrix-core form against the FP64 one for every particle
    # (I(q) pointwise where I >= 1e-3 max I)
    q = st.q_mesh(0.2, 25.0, 2048)
    qd = torch.from_numpy(q).to(dev)
    S = torch.empty(b.ntypes_total * q.size, dtype=torch.float64, device=dev)
    Sm = torch.empty_like(S)
    b.run_dev(hist.data_ptr(), qd.data_ptr(), q.size, S.data_ptr())
    b.run_dev(hist.data_ptr(), qd.data_ptr(), q.size, Sm.data_ptr(), mfma=True)
    bctx.synchronize()
    S, Sm = S.cpu().numpy(), Sm.cpu().numpy()
    worst = 0.0
    for s, p in enumerate(parts):
        P = p.n_el * (p.n_el + 1) // 2
        t0 = int(b.type_first[s])
        sl = np.linspace(2.0, 9.0, p.n_el)
        Ss = S[t0 * q.size:(t0 + P) * q.size].reshape(P, q.size)
        Sms = Sm[t0 * q.size:(t0 + P) * q.size].reshape(P, q.size)
        If = O.intensity_1d(Ss, p.el_count, q, sl=sl)
        worst = max(worst, rel_err(O.intensity_1d(Sms, p.el_count, q, sl=sl), If)[0])
        if s in sample:
            nh = int(b.nhist[s])
            h = H[int(b.hist_off[s]):int(b.hist_off[s]) + P * nh].reshape(P, nh)
            pw, l2 = rel_err(If, O.intensity_1d(O.hist_pairsum(h, BIN, q), p.el_count, q, sl=sl))
            assert pw <= TOL and l2 <= TOL, s
    print(f"full ensemble: matrix-core I(q) within {worst:.2e} of the FP64 pair sums", flush=True)
    assert worst <= TOL
    b.close()


def test_batch_mfma_pairsums(bctx, ensemble):
    """XS_PAIRSUM_MFMA in a batch run: the pair types grouped by count slices
    (1-4 here, from pair counts of 1 to 1.8e7) and bin count, on the int8
    matrix cores.  S within 1e-9 of max|S| per structure and I(q) within 1e-5
    of the FP64 oracle; histograms and PDFs as without the flag; runs with
    and without the flag alternate on one plan; graph replays on the device
    reproduce the host-mode run bit for bit."""
    import torch
    import xansons_hip as xh
    q = st.q_mesh(0.5, 20.0, 600)
    sl, rho = sl_for(ensemble), rho_for(ensemble)
    b = xh.Batch(bctx, [(e.xyz64(), e.el_count) for e in ensemble], BIN, f64=True)
    h_v, S_v, p_v = b.run(q=q, pdf_type="rPDF", rho0=rho, sl=sl)
    h_m, S_m, p_m = b.run(q=q, pdf_type="rPDF", rho0=rho, sl=sl, mfma=True)
    assert np.array_equal(h_m, h_v) and np.array_equal(p_m, p_v)
    so = 0
    for s, (e, (h, Ss, _)) in enumerate(zip(ensemble, split(b, ensemble, h_m, S_m, p_m, q.size))):
        sle = sl[so:so + e.n_el]
        so += e.n_el
        So = O.hist_pairsum(h, BIN, q)
        srel = float(np.max(np.abs(Ss - So)) / max(np.abs(So).max(), 1e-300))
        Io = O.intensity_1d(So, e.el_count, q, sl=sle)
        pw, l2 = rel_err(O.intensity_1d(Ss, e.el_count, q, sl=sle), Io)
        print(f"struct {s}: N={e.n} S {srel:.2e} I {pw:.2e} {l2:.2e}", flush=True)
        assert srel <= 1e-9 and pw <= TOL and l2 <= TOL, s
    # without the flag again (its chunk sums meet in FP64 atomics, so to
    # rounding), then with it: the same results
    _, S_v2, _ = b.run(q=q)
    assert np.allclose(S_v2, S_v, rtol=1e-12, atol=1e-12 * np.abs(S_v).max())
    dev = torch.device("cuda", 0)
    qd = torch.from_numpy(q).to(dev)
    hist = torch.empty(b.hist_len, dtype=torch.int64, device=dev)
    S = torch.empty(b.ntypes_total * q.size, dtype=torch.float64, device=dev)
    for rep in range(3):
        S.fill_(float("nan"))
        torch.cuda.synchronize()
        b.run_dev(hist.data_ptr(), qd.data_ptr(), q.size, S.data_ptr(), graph=True, mfma=True)
        bctx.synchronize()
        assert np.array_equal(S.cpu().numpy(), S_m), rep
        print(f"mfma replay {rep}: ok", flush=True)
    b.close()


def test_batch_mfma_refuses_2p35_pairs(bctx):
    """A pair type of 2^35 pairs or more needs more count slices than the
    kernels take: the run is refused before any launch (checked on the
    plan's pair counts, the structure is never binned)."""
    import xansons_hip as xh
    # 262,145 atoms of one element: 3.4e10 pairs >= 2^35
    n = 262145
    x = np.zeros((3, n))
    x[0] = np.arange(n) * 1e-3
    b = xh.Batch(bctx, [(x, [n])], 1.0, f64=True)
    with pytest.raises(RuntimeError, match="2\\^35"):
        b.run(q=st.q_mesh(0.5, 5.0, 8), mfma=True)
    b.close()
