"""Pins of the oracle that the reference's printed outputs leave open:
closed forms, invariants and a brute-force cross-check of the FP32 binning
(the GPU expression, whose reference builds are fast-math and cannot be
built here)."""
import itertools
import math

import numpy as np
import pytest

import xs_oracle as O
from xansons_hip import structures as st


def brute_f32(xyz32, el_count, bin_, nhist):
    """Independent numpy float32 evaluation of (unsigned)(sqrtf((dx^2+dy^2)+dz^2)/bin)."""
    xyz32 = np.asarray(xyz32, np.float32)
    offs = np.concatenate([[0], np.cumsum(el_count)])
    out = []
    b = np.float32(bin_)
    for a in range(len(el_count)):
        for c in range(a, len(el_count)):
            A = xyz32[:, offs[a]:offs[a + 1]]
            B = xyz32[:, offs[c]:offs[c + 1]]
            d = [A[k][:, None] - B[k][None, :] for k in range(3)]
            s = d[0] * d[0] + d[1] * d[1]
            s = s + d[2] * d[2]
            k = np.trunc(np.sqrt(s) / b).astype(np.int64)
            if a == c:
                k = k[np.triu_indices(A.shape[1], 1)]
            out.append(np.bincount(k.ravel(), minlength=nhist)[:nhist])
    return np.array(out, dtype=np.uint64)


def test_pair_index_order():
    n = 4
    got = [O.lib().or_pair_index(a, b, n) for a in range(n) for b in range(a, n)]
    assert got == list(range(10))


def test_hist_f32_matches_numpy_float32():
    ens = st.random_cluster([60, 45, 30], box=20.0, min_dist=0.8, seed=5)
    nh = O.hist_size(ens.xyz64(), 0.01)
    h, ov = O.hist_f32(ens.xyz, ens.el_count, 0.01, nh)
    assert ov == 0
    assert np.array_equal(h, brute_f32(ens.xyz, ens.el_count, 0.01, nh))


def test_hist_f32_differs_from_f64_only_at_bin_edges():
    """The FP32 and FP64 bin indices agree except for pairs whose distance
    lies within float rounding of a bin edge; a wrong expression (rounding
    instead of truncation, a dropped term) breaks this immediately."""
    ens = st.fcc_sphere(radius=14.0, seed=9)
    x64 = ens.xyz64()
    b = 0.01
    nh = O.hist_size(x64, b)
    h32, _ = O.hist_f32(ens.xyz, ens.el_count, b, nh)
    h64, _ = O.hist_f64(x64, ens.el_count, b, nh)
    assert h32.sum() == h64.sum()
    diff = h32.astype(np.int64) - h64.astype(np.int64)
    moved = np.abs(diff).sum() // 2
    # count pairs whose exact distance is within 4 float ulps of an edge
    d = np.sqrt(sum((x64[k][:, None] - x64[k][None, :]) ** 2 for k in range(3)))
    d = d[np.triu_indices(ens.n, 1)] / b
    near = np.abs(d - np.rint(d)) <= 4 * np.spacing(np.float32(d.max())) * 1.0
    assert moved <= near.sum()
    assert moved < 1e-3 * h64.sum()
    # every change moves a count to the neighbouring bin only
    assert np.abs(np.convolve(diff[0], np.ones(2, dtype=np.int64), "valid")).max() <= moved


def test_hist_simple_cubic_closed_form():
    """Perfect n^3 simple-cubic lattice: offset (i,j,k) occurs
    (n-|i|)(n-|j|)(n-|k|) times as an ordered pair; shells never sit on an edge."""
    n, a, b = 7, 1.0, 0.0973
    g = np.arange(n, dtype=np.float64) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    nh = O.hist_size(xyz, b)
    expect = np.zeros(nh, dtype=np.uint64)
    for i, j, k in itertools.product(range(-(n - 1), n), repeat=3):
        if (i, j, k) == (0, 0, 0):
            continue
        r = math.sqrt(i * i + j * j + k * k) * a / b
        assert abs(r - round(r)) > 1e-6
        expect[int(r)] += (n - abs(i)) * (n - abs(j)) * (n - abs(k))
    expect //= 2
    h64, _ = O.hist_f64(xyz, [n ** 3], b, nh)
    h32, _ = O.hist_f32(xyz.astype(np.float32), [n ** 3], b, nh)
    assert np.array_equal(h64[0], expect)
    assert np.array_equal(h32[0], expect)


def test_hist_permutation_invariance_and_totals():
    ens = st.random_cluster([70, 50], box=18.0, min_dist=0.9, seed=6)
    nh = O.hist_size(ens.xyz64(), 0.005)
    h1, _ = O.hist_f32(ens.xyz, ens.el_count, 0.005, nh)
    rng = np.random.default_rng(0)
    perm = np.concatenate([rng.permutation(70), 70 + rng.permutation(50)])
    h2, _ = O.hist_f32(ens.xyz[:, perm], ens.el_count, 0.005, nh)
    assert np.array_equal(h1, h2)
    assert [int(v) for v in h1.sum(axis=1)] == [70 * 69 // 2, 70 * 50, 50 * 49 // 2]


def test_hist_overflow_counted():
    xyz = np.array([[0.0, 10.0], [0.0, 0.0], [0.0, 0.0]])
    h, ov = O.hist_f64(xyz, [2], 1.0, 5)
    assert h.sum() == 0 and ov == 1


def test_hist_size_closed_form():
    xyz = np.array([[0.0, 3.0], [0.0, 4.0], [0.0, 0.0]])
    assert O.hist_size(xyz, 0.5) == 11            # |(3,4,0)| = 5 -> 5/0.5 + 1


def test_debye_two_atoms_closed_form():
    d = 2.345
    xyz = np.array([[0.0, d], [0.0, 0.0], [0.0, 0.0]])
    q = np.linspace(0.0, 20.0, 101)
    S = O.debye_pairsum(xyz, [2], q)
    np.testing.assert_allclose(S[0], np.sin(q * d) / (q * d + 1e-8), rtol=1e-14, atol=1e-15)
    I = O.intensity_1d(S, [2], q, sl=np.array([3.0]))
    np.testing.assert_allclose(I, (2 * 9.0 + 2 * 9.0 * S[0]) / 2, rtol=1e-14)


def test_hist_pairsum_equals_debye_on_bin_centres():
    """With the atom pair at a bin centre the histogram sum is the direct
    Debye sum."""
    b = 0.25
    q = np.linspace(0.3, 15.0, 64)
    for k in (2, 8, 21):
        xyz = np.array([[0.0, (k + 0.5) * b], [0.0, 0.0], [0.0, 0.0]])
        h, _ = O.hist_f64(xyz, [2], b, O.hist_size(xyz, b))
        assert h[0, k] == 1
        np.testing.assert_allclose(O.hist_pairsum(h, b, q), O.debye_pairsum(xyz, [2], q),
                                   rtol=1e-12, atol=1e-13)


def test_polar_factor_closed_form():
    lam = 1.5
    q = np.array([0.0, 4 * np.pi / lam * np.sin(np.pi / 4)])   # 2 theta = 0 and 90 deg
    I = O.polar_1d(np.ones(2), q, lam)
    np.testing.assert_allclose(I, [1.0, 0.5], atol=1e-15)


def test_rotations_orthonormal():
    rng = np.random.default_rng(1)
    for conv in range(12):
        R = O.rot_matrix(rng.uniform(-3, 3, 3), conv)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1) < 1e-13


def test_amp2d_closed_forms():
    ens = st.random_cluster([7, 5], box=8.0, min_dist=1.0, seed=3)
    x = ens.xyz64()
    rng = np.random.default_rng(2)
    qv = rng.normal(size=(20, 3))
    qv[0] = 0
    rot = np.stack([O.rot_matrix(rng.uniform(-3, 3, 3), 5).ravel() for _ in range(3)])
    ff = rng.uniform(0.5, 2.0, size=(2, 20))
    I = O.amp2d(x, ens.el_count, qv, rot, ff)
    # q = 0: |sum f|^2 per rotation
    assert abs(I[0] - 3 * (7 * ff[0, 0] + 5 * ff[1, 0]) ** 2) < 1e-9
    # Friedel: I(q) = I(-q)
    np.testing.assert_allclose(O.amp2d(x, ens.el_count, -qv, rot, ff), I, rtol=1e-12)
    # rotating q by R^T equals rotating the atoms by R
    one = np.eye(3).ravel()[None]
    for r in rot:
        R = r.reshape(3, 3)
        a = O.amp2d(x, ens.el_count, qv, r[None], ff)
        b = O.amp2d(R @ x, ens.el_count, qv, one, ff)
        np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-10)
    # one atom: f^2
    single = O.amp2d(np.zeros((3, 1)), [1], qv, one, ff[:1])
    np.testing.assert_allclose(single, ff[0] ** 2, rtol=1e-14)


def test_orientation_average_equals_debye():
    """Mean of |A|^2 over uniformly random orientations is the Debye intensity
    (statistical, 4000 rotations)."""
    ens = st.random_cluster([6], box=6.0, min_dist=1.2, seed=4)
    x = ens.xyz64()
    rng = np.random.default_rng(7)
    qs = np.array([0.7, 1.9, 3.1])
    qv = np.stack([qs, 0 * qs, 0 * qs], axis=1)
    u = rng.normal(size=(4000, 4))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    w, xq, yq, zq = u.T
    R = np.stack([1 - 2 * (yq * yq + zq * zq), 2 * (xq * yq - zq * w), 2 * (xq * zq + yq * w),
                  2 * (xq * yq + zq * w), 1 - 2 * (xq * xq + zq * zq), 2 * (yq * zq - xq * w),
                  2 * (xq * zq - yq * w), 2 * (yq * zq + xq * w), 1 - 2 * (xq * xq + yq * yq)], axis=1)
    I = O.amp2d(x, [6], qv, R, np.ones((1, 3))) / 4000
    S = O.debye_pairsum(x, [6], qs)
    debye = 6 + 2 * S[0]
    np.testing.assert_allclose(I, debye, rtol=0.05)


def test_density_uniform_block():
    """A filled simple-cubic block with spacing a has density 1/a^3 inside."""
    a = 1.1
    g = np.arange(31) * a
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    assert abs(O.atomic_density(xyz) * a ** 3 - 1) < 0.05
