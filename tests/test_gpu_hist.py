"""H1 parity: the HIP histogram equals the oracle bit for bit (FP32 binning
against or_hist_f32, FP64 binning against or_hist_f64, which matches the
reference CPU executable)."""
import numpy as np
import pytest
try:
    import torch  # noqa: F401  (first import can take a minute on a fresh box)
except ImportError:
    torch = None

import xs_oracle as O
from conftest import golden
from xansons_hip import structures as st

pytestmark = pytest.mark.gpu


def both_modes(ctx, ens, bin_, nhist=None):
    x64 = ens.xyz64()
    nh = nhist or O.hist_size(x64, bin_)
    g32, ov32 = ctx.hist(ens.xyz, ens.el_count, bin_, nh, f64=False)
    o32, oo32 = O.hist_f32(ens.xyz, ens.el_count, bin_, nh)
    g64, ov64 = ctx.hist(x64, ens.el_count, bin_, nh, f64=True)
    o64, oo64 = O.hist_f64(x64, ens.el_count, bin_, nh)
    return (g32, ov32, o32, oo32), (g64, ov64, o64, oo64)


def assert_same(pair, label):
    g, ov, o, oo = pair
    print(f"{label}: total={int(g.sum())} mismatched_bins={int(np.sum(g != o))} ovf={ov}/{oo}", flush=True)
    assert ov == oo
    assert np.array_equal(g, o)


def test_fcc_sphere_c1(ctx):
    g = golden("fcc1985")
    ens = st.Ensemble(g["xyz"], list(g["el_count"]), ["Au"])
    r32, r64 = both_modes(ctx, ens, 0.01)
    assert_same(r32, "c1 f32")
    assert_same(r64, "c1 f64")
    # FP64 binning also equals the reference executable's own counts
    rec = np.rint(g["rdf"] * 0.01 * 1985 / 2).astype(np.uint64)
    assert np.array_equal(r64[0][0, :rec.size], rec)


def test_rocksalt_small_three_types(ctx):
    g = golden("nacl_small")
    ens = st.Ensemble(g["xyz"], list(g["el_count"]), ["Na", "Cl"])
    r32, r64 = both_modes(ctx, ens, 0.005)
    assert_same(r32, "nacl f32")
    assert_same(r64, "nacl f64")


@pytest.mark.parametrize("counts", [[1], [2], [1023], [1024, 1], [1025, 0, 3], [2049, 700],
                                    [5, 6, 7, 3000]])
def test_ragged_sizes(ctx, counts):
    ens = st.random_cluster([max(c, 0) for c in counts], box=8.0 + 0.02 * sum(counts) ** (1 / 3) * 10,
                            min_dist=0.05, seed=sum(counts)) if sum(counts) <= 4000 else None
    r32, r64 = both_modes(ctx, ens, 0.01)
    assert_same(r32, f"ragged {counts} f32")
    assert_same(r64, f"ragged {counts} f64")


def test_sixteen_elements_and_seventeen_rejected(ctx):
    """The most elements a structure may have (16, 136 pair types), with
    empty and one-atom elements among them, both binnings bit-exact; a 17th
    element is rejected with XS_ERR_RANGE by the single-structure call and
    by a batch plan, before anything runs."""
    import xansons_hip as xh
    counts = [37, 0, 64, 1, 129, 5, 300, 2, 0, 77, 600, 3, 9, 200, 1, 50]
    ens = st.random_cluster(counts, box=16.0, min_dist=0.3, seed=1616)
    r32, r64 = both_modes(ctx, ens, 0.01)
    assert r64[0].shape[0] == 136
    assert_same(r32, "16 elements f32")
    assert_same(r64, "16 elements f64")
    el17 = counts + [4]
    x17 = np.ascontiguousarray(np.concatenate([ens.xyz64(), np.full((3, 4), 9.0) + np.arange(4)], axis=1))
    for f64 in (False, True):
        x = x17 if f64 else x17.astype(np.float32)
        with pytest.raises(xh.XSError) as e:
            ctx.hist(x, el17, 0.01, 4000, f64=f64)
        assert e.value.code == -5
        with pytest.raises(xh.XSError) as e:
            xh.Batch(ctx, [(ens.xyz64(), counts), (x17, el17)], 0.01, f64=f64)
        assert e.value.code == -5


def test_sliding_window_large_nhist(ctx):
    """nhist far above the LDS window: the window slides and flushes."""
    ens = st.rocksalt_cube(ncell=18, vacancy=0.02, sigma=0.08, seed_vac=31, seed_jit=32)
    r32, r64 = both_modes(ctx, ens, 0.001)
    assert r32[0].shape[1] > 100_000
    assert_same(r32, "window f32")
    assert_same(r64, "window f64")


def test_direct_global_path_tiny_bin(ctx):
    """Bins so fine that one tile pair spans more than the LDS window: those
    pairs take global atomics directly."""
    ens = st.random_cluster([1500, 1200], box=40.0, min_dist=0.5, seed=77)
    r32, r64 = both_modes(ctx, ens, 0.0002)
    assert_same(r32, "direct f32")
    assert_same(r64, "direct f64")


def test_overflow_when_nhist_too_small(ctx):
    ens = st.fcc_sphere(radius=9.0, seed=3)
    r32, r64 = both_modes(ctx, ens, 0.01, nhist=700)
    assert r32[1] > 0
    assert_same(r32, "overflow f32")
    assert_same(r64, "overflow f64")


def test_bin_edges_integer_lattice(ctx):
    """Integer-lattice atoms with bins of 0.5, 0.25 and 0.1 A: most distances
    (1, 2, 3, sqrt(k)...) fall exactly on or within rounding of bin edges, so
    the FP64 fast path must hand them to the exact IEEE evaluation."""
    r = np.arange(12, dtype=np.float64)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=0).reshape(3, -1)
    ens = st.Ensemble(np.ascontiguousarray(g.astype(np.float32)), [g.shape[1] // 2, g.shape[1] - g.shape[1] // 2],
                      ["A", "B"])
    for b in (0.5, 0.25, 0.1):
        r32, r64 = both_modes(ctx, ens, b)
        assert_same(r64, f"lattice bin {b} f64")
        assert_same(r32, f"lattice bin {b} f32")


def test_f32_root_rounding_one_ulp_bins(ctx):
    """FP32 binning with bin = 2^-23 A and all distances below 2 A: every
    distance in [1, 2) A then has a bin of its own float value (r / bin is an
    exact integer there), so the counts record each pair's rounded sqrtf(s) bit
    for bit: 4.2 million cross-cluster pairs, plus near-coincident atoms with
    s down in the denormal range, whose bin is 0."""
    rng = np.random.Generator(np.random.PCG64(21))
    a = rng.uniform(-0.15, 0.15, size=(3, 2040))
    b = rng.uniform(-0.15, 0.15, size=(3, 2048)) + np.array([[1.15], [0.3], [0.1]])
    tiny = np.array([[0.0, 1e-20, 3e-21, 1e-19, -2e-22, 5e-24, 1e-18, 0.0],
                     [0.0, 0.0, 2e-21, 0.0, 1e-22, 0.0, -1e-18, 4e-23],
                     [0.0, 0.0, 0.0, 3e-20, 0.0, 5e-24, 0.0, 0.0]])
    xyz = np.ascontiguousarray(np.concatenate([tiny, a, b], axis=1).astype(np.float32))
    ens = st.Ensemble(xyz, [xyz.shape[1]], ["A"])
    bin_, nh = 2.0 ** -23, 1 << 24
    g, ov = ctx.hist(ens.xyz, ens.el_count, bin_, nh, f64=False)
    o, oo = O.hist_f32(ens.xyz, ens.el_count, bin_, nh)
    ulp_bins = int(np.count_nonzero(o[..., 1 << 23:]))
    print(f"distinct roots in [1, 2): {ulp_bins}", flush=True)
    assert ulp_bins > 1_000_000
    assert_same((g, ov, o, oo), "one-ulp bins f32")


@pytest.mark.parametrize("double_quotient", [False, True])
def test_f32_quotient_forms(ctx, double_quotient):
    """FP32 binning with the FP32 quotient form (taken when the bin passes
    the device's check over all significands) and with the double product
    (XS_F32_DOUBLE_QUOTIENT): both equal the oracle bit for bit on bin-edge
    distances (integer lattice) and a jittered rock-salt block with
    vacancies on the host path, and on a sphere on the device path."""
    import torch
    r = np.arange(12, dtype=np.float64)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=0).reshape(3, -1)
    lat = st.Ensemble(np.ascontiguousarray(g.astype(np.float32)), [g.shape[1] // 2, g.shape[1] - g.shape[1] // 2],
                      ["A", "B"])
    cl = st.rocksalt_cube(ncell=8, vacancy=0.05, sigma=0.1, seed_vac=41, seed_jit=42)
    for ens, b in ((lat, 0.5), (lat, 0.1), (cl, 0.0123), (cl, 1.0 / 3.0)):
        nh = O.hist_size(ens.xyz64(), b)
        gh, gov = ctx.hist(ens.xyz, ens.el_count, b, nh, double_quotient=double_quotient)
        assert_same((gh, gov, *O.hist_f32(ens.xyz, ens.el_count, b, nh)), f"quotient {double_quotient} bin {b}")
    ens = st.fcc_sphere(radius=22.0, seed=6)
    nh = O.hist_size(ens.xyz64(), 0.004)
    dev = torch.device("cuda", 0)
    x = torch.from_numpy(np.ascontiguousarray(ens.xyz)).to(dev)
    h = torch.empty(nh, dtype=torch.int64, device=dev)
    ov = torch.empty(1, dtype=torch.int64, device=dev)
    ctx.hist_dev(x.data_ptr(), ens.n, ens.el_count, 0.004, nh, h.data_ptr(), ov.data_ptr(),
                 double_quotient=double_quotient)
    ctx.synchronize()
    ho, oo = O.hist_f32(ens.xyz, ens.el_count, 0.004, nh)
    assert np.array_equal(h.cpu().numpy().astype(np.uint64), ho[0])
    assert int(ov.item()) == oo


def test_device_pointers_match_host(ctx):
    import torch
    ens = st.fcc_sphere(radius=12.0, seed=4)
    nh = O.hist_size(ens.xyz64(), 0.01)
    host, _ = ctx.hist(ens.xyz, ens.el_count, 0.01, nh)
    dev = torch.device("cuda", 0)
    x = torch.from_numpy(np.ascontiguousarray(ens.xyz)).to(dev)
    h = torch.empty(nh, dtype=torch.int64, device=dev)
    ov = torch.empty(1, dtype=torch.int64, device=dev)
    ctx.hist_dev(x.data_ptr(), ens.n, ens.el_count, 0.01, nh, h.data_ptr(), ov.data_ptr())
    ctx.synchronize()
    assert np.array_equal(h.cpu().numpy().astype(np.uint64), host[0])
    assert int(ov.item()) == 0


def test_rocksalt_million_properties(ctx):
    """Full 980,219-atom rock-salt cube: exact totals, no overflow, repeatable,
    FP32/FP64 agree to bin-edge pairs, and a 24k-atom subset is bit-exact."""
    ens = st.rocksalt_cube()
    x64 = ens.xyz64()
    nh = O.hist_size(x64, 0.005)
    h1, ov1 = ctx.hist(ens.xyz, ens.el_count, 0.005, nh)
    h2, ov2 = ctx.hist(ens.xyz, ens.el_count, 0.005, nh)
    na, nb = ens.el_count
    assert ov1 == 0 and ov2 == 0
    assert [int(v) for v in h1.sum(axis=1)] == [na * (na - 1) // 2, na * nb, nb * (nb - 1) // 2]
    assert np.array_equal(h1, h2)
    h64, ov64 = ctx.hist(x64, ens.el_count, 0.005, nh, f64=True)
    moved = np.abs(h1.astype(np.int64) - h64.astype(np.int64)).sum() // 2
    print(f"c2 f32 vs f64 moved pairs {moved} of {int(h1.sum())}", flush=True)
    assert moved < 1e-3 * h1.sum()
    sub = st.Ensemble(np.ascontiguousarray(np.concatenate(
        [ens.element(0)[:, ::40], ens.element(1)[:, ::40]], axis=1)),
        [ens.element(0)[:, ::40].shape[1], ens.element(1)[:, ::40].shape[1]], ens.names)
    r32, r64 = both_modes(ctx, sub, 0.005, nhist=nh)
    assert_same(r32, "c2 subset f32")
    assert_same(r64, "c2 subset f64")


@pytest.mark.parametrize("shape,n,n_el,bin_", [("fcc", 30000, 3, 0.01), ("core-shell", 20000, 2, 0.005),
                                              ("cube", 25000, 1, 0.01), ("bcc", 12000, 2, 0.002)])
def test_device_path_equal_count_tiles(ctx, shape, n, n_el, bin_):
    """Device-pointer path (equal-count tiles by stable radix sorts, or the
    cell grid when a tile would exceed the window) equals the oracle, FP64
    and FP32 binning; the bcc case has more bins than the largest window."""
    import torch
    rng = np.random.Generator(np.random.PCG64(n))
    ens = st.particle(n, n_el, shape, 3.9 if shape != "bcc" else 3.1, 0.05, rng)
    x64 = ens.xyz64()
    nh = O.hist_size(x64, bin_)
    dev = torch.device("cuda", 0)
    P = ens.n_el * (ens.n_el + 1) // 2
    for f64 in (True, False):
        xs = x64 if f64 else ens.xyz
        x = torch.from_numpy(np.ascontiguousarray(xs)).to(dev)
        h = torch.empty(P * nh, dtype=torch.int64, device=dev)
        ov = torch.empty(1, dtype=torch.int64, device=dev)
        ctx.hist_dev(x.data_ptr(), ens.n, ens.el_count, bin_, nh, h.data_ptr(), ov.data_ptr(),
                     f64=f64)
        ctx.synchronize()
        ho, oo = (O.hist_f64(x64, ens.el_count, bin_, nh) if f64 else
                  O.hist_f32(ens.xyz, ens.el_count, bin_, nh))
        assert np.array_equal(h.cpu().numpy().astype(np.uint64).reshape(ho.shape), ho), (shape, f64)
        assert int(ov.item()) == oo


def _hist64_host_and_device(ctx, x64, el_count, bin_, nh):
    import torch
    host, ov_h = ctx.hist(x64, el_count, bin_, nh, f64=True)
    dev = torch.device("cuda", 0)
    x = torch.from_numpy(np.ascontiguousarray(x64)).to(dev)
    P = len(el_count) * (len(el_count) + 1) // 2
    h = torch.empty(P * nh, dtype=torch.int64, device=dev)
    ov = torch.empty(1, dtype=torch.int64, device=dev)
    ctx.hist_dev(x.data_ptr(), x64.shape[1], el_count, bin_, nh, h.data_ptr(), ov.data_ptr(), f64=True)
    ctx.synchronize()
    return host, ov_h, h.cpu().numpy().astype(np.uint64).reshape(host.shape), int(ov.item())


@pytest.mark.parametrize("case", ["far_from_origin", "sparse_tiles", "shifted_integer_lattice"])
def test_fp64_tile_centred_distances(ctx, case):
    """The FP64 fast path forms s = |ri'|^2 + |rj'|^2 - 2 ri'.rj' from
    coordinates relative to the I tile's centre; its band must send every
    pair whose estimate could fall in another bin to the exact evaluation.
    Cases: FP64 coordinates 10^4 A from the origin (rounding of ri - c),
    a sparse cluster whose tiles span the whole structure (large tile
    radius), and an integer lattice shifted by 1000.25 A (exact bin edges
    with rounded, shifted coordinates).  Host and device paths, bit-exact."""
    rng = np.random.Generator(np.random.PCG64(2718))
    if case == "far_from_origin":
        ens = st.fcc_sphere(radius=24.0, seed=5)
        x64 = ens.xyz64() + rng.normal(0, 0.05, ens.xyz.shape) + np.array([[1.0e4], [-3.1e4], [5.3e3]])
        el, bin_ = ens.el_count, 0.01
    elif case == "sparse_tiles":
        x64 = rng.uniform(-40.0, 40.0, (3, 2500))
        el, bin_ = [1300, 1200], 0.02
    else:
        r = np.arange(11, dtype=np.float64)
        g = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=0).reshape(3, -1)
        x64 = g + 1000.25
        el, bin_ = [g.shape[1] // 2, g.shape[1] - g.shape[1] // 2], 0.25
    x64 = np.ascontiguousarray(x64)
    nh = O.hist_size(x64, bin_)
    want, oo = O.hist_f64(x64, el, bin_, nh)
    host, ov_h, devh, ov_d = _hist64_host_and_device(ctx, x64, el, bin_, nh)
    print(f"{case}: nhist={nh} pairs={int(want.sum())} host mismatched={int(np.sum(host != want))} "
          f"device mismatched={int(np.sum(devh != want))}", flush=True)
    assert ov_h == oo and ov_d == oo
    assert np.array_equal(host, want)
    assert np.array_equal(devh, want)


@pytest.mark.parametrize("nhist", [1 << 19, (1 << 19) + 1, None])
def test_fp64_grid_at_the_fixed_point_limit(ctx, nhist):
    """Grids of at most 2^19 bins take the FP64 fast path in fixed point (the
    estimate's floor and fraction read from the words of estimate + 2^20),
    wider ones the fract form.  A random cluster whose distances span about
    495,000 bins of 1.4e-4 A, binned on 2^19 and 2^19 + 1 bins, and on its
    own 693,000 bins of 1e-4 A: host and device paths, bit-exact."""
    rng = np.random.Generator(np.random.PCG64(1931))
    x64 = np.ascontiguousarray(rng.uniform(-20.0, 20.0, (3, 2500)))
    el = [1400, 1100]
    bin_ = 1.4e-4 if nhist else 1.0e-4
    nh = nhist or O.hist_size(x64, bin_)
    assert nh >= O.hist_size(x64, bin_)
    want, oo = O.hist_f64(x64, el, bin_, nh)
    host, ov_h, devh, ov_d = _hist64_host_and_device(ctx, x64, el, bin_, nh)
    print(f"nhist={nh} pairs={int(want.sum())} last bin={int(np.nonzero(want.sum(axis=0))[0].max())} "
          f"host mismatched={int(np.sum(host != want))} device mismatched={int(np.sum(devh != want))}",
          flush=True)
    assert ov_h == oo and ov_d == oo
    assert np.array_equal(host, want)
    assert np.array_equal(devh, want)


def test_fp64_distances_beyond_2p20_bins_on_a_small_grid(ctx):
    """A given grid of 200,000 bins (the fixed-point fast path) with
    distances of up to 1.9 million bins: estimates at or above 2^20 bins leave
    the fixed-point range of the fast path; those pairs lie past the grid and
    must all reach the overflow counter, the rest their bins.  Host and
    device paths, bit-exact, overflow equal."""
    rng = np.random.Generator(np.random.PCG64(4099))
    x64 = np.ascontiguousarray(rng.uniform(-16.0, 16.0, (3, 1500)))
    el, bin_, nh = [800, 700], 3e-5, 200000
    assert O.hist_size(x64, bin_) > (1 << 20)
    want, oo = O.hist_f64(x64, el, bin_, nh)
    host, ov_h, devh, ov_d = _hist64_host_and_device(ctx, x64, el, bin_, nh)
    print(f"overflow={oo} in-grid={int(want.sum())} host mismatched={int(np.sum(host != want))} "
          f"device mismatched={int(np.sum(devh != want))} ovf {ov_h}/{ov_d}", flush=True)
    assert oo > 0
    assert ov_h == oo and ov_d == oo
    assert np.array_equal(host, want)
    assert np.array_equal(devh, want)


def test_fp64_coincident_atoms_of_two_elements(ctx):
    """Two elements on the same sites (a mixed site written out as two atoms):
    the cross-type pairs lie at distance 0, where the tile-centred squared
    distance rounds to 0 or a tiny value of either sign and the fast path's
    estimate is NaN, infinite or near 0.  They must land in bin 0 as in the
    reference expression; FP64 and FP32 binning, bit-exact."""
    ens = st.fcc_sphere(radius=14.0, seed=9)
    x64 = ens.xyz64()
    n = x64.shape[1]
    both = st.Ensemble(np.ascontiguousarray(np.concatenate([ens.xyz, ens.xyz], axis=1)), [n, n],
                       ["A", "B"])
    r32, r64 = both_modes(ctx, both, 0.01)
    assert r64[2][1][0] >= n
    assert_same(r64, "coincident f64")
    assert_same(r32, "coincident f32")


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_nan_coordinate_rejected(ctx, f64, axis):
    """A NaN anywhere (not only in the first atom) is rejected with
    XS_ERR_ARG before any pair kernel runs: host input, device input and
    batch plans.  fmin/fmax and std::min/max would drop it from the
    bounding box and let a NaN distance reach an unchecked LDS slot."""
    import xansons_hip as xh
    ens = st.random_cluster([300, 200], box=14.0, min_dist=0.5, seed=5)
    xyz = ens.xyz64().copy()
    xyz[axis, 337] = np.nan
    nh = 4000
    print(f"nan test f64={f64} axis={axis}", flush=True)
    with pytest.raises(xh.XSError) as e:
        ctx.hist(xyz, ens.el_count, 0.01, nh, f64=f64)
    assert e.value.code == -1 and "non-finite" in str(e.value)
    dev = torch.device("cuda", 0)
    xd = torch.from_numpy(np.ascontiguousarray(xyz if f64 else xyz.astype(np.float32))).to(dev)
    h = torch.zeros(3 * nh, dtype=torch.int64, device=dev)
    with pytest.raises(xh.XSError) as e:
        ctx.hist_dev(xd.data_ptr(), ens.n, ens.el_count, 0.01, nh, h.data_ptr(), f64=f64)
    assert e.value.code == -1 and "non-finite" in str(e.value)
    ok = ens.xyz64()
    with pytest.raises(xh.XSError) as e:
        xh.Batch(ctx, [(ok, ens.el_count), (xyz, ens.el_count), (ok, ens.el_count)], 0.01, f64=f64)
    assert e.value.code == -1 and "non-finite" in str(e.value)
    # the same structure without the NaN still runs and equals the oracle
    g, _ = ctx.hist(ok, ens.el_count, 0.01, nh, f64=True)
    o, _ = O.hist_f64(ok, ens.el_count, 0.01, nh)
    assert np.array_equal(g, o)


def test_f32_binning_overflowing_coordinates_rejected(ctx):
    """Coordinates beyond the float range become infinite when a batch plan
    (double input) rounds them to float for FP32 binning: rejected as
    non-finite.  The bin of 1e31 A keeps the double span itself within 2^31
    bins and the bin count is given, so it is the float check that refuses."""
    import xansons_hip as xh
    ens = st.random_cluster([300], box=10.0, min_dist=0.5, seed=6)
    xyz = ens.xyz64().copy()
    xyz[0, 17] = 1e39
    with pytest.raises(xh.XSError) as e:
        xh.Batch(ctx, [(xyz, ens.el_count)], 1e31, nhist=[1000], f64=False)
    assert e.value.code == -1 and "non-finite" in str(e.value)
