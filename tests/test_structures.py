"""Synthetic workload generators: seeded, element-major, within their stated bounds."""
import numpy as np

from xansons_hip import structures as st


def test_uniform_rotations_are_rotations_and_spread():
    R = st.uniform_rotations(4096, seed=8)
    assert np.abs(np.einsum("nij,nkj->nik", R, R) - np.eye(3)).max() < 1e-14
    assert np.allclose(np.linalg.det(R), 1.0, atol=1e-14)
    # uniform on SO(3): E[R] = 0, E[trace R] = 0
    assert np.abs(R.mean(axis=0)).max() < 0.05
    assert abs(np.trace(R, axis1=1, axis2=2).mean()) < 0.05
    assert np.array_equal(R, st.uniform_rotations(4096, seed=8))


def test_q_plane_symmetric():
    q = st.q_plane(8.0, 64)
    assert q.shape == (4096, 3)
    assert np.array_equal(q[::-1, :2], -q[:, :2])


def test_particle_ensemble_bounds():
    ens = st.particle_ensemble(40, seed=3, nmin=300, nmax=5000)
    for e in ens:
        assert 1 <= e.n_el <= 4 and all(c > 0 for c in e.el_count)
        assert 150 <= e.n <= 6000
        assert e.xyz.dtype == np.float32 and e.xyz.shape == (3, e.n)


def test_c4_particles_fit_16384_bins_at_0_004():
    for p in st.c4_particles(24, seed=11):
        x = p.xyz64()
        r = np.sqrt(((x - x.mean(axis=1, keepdims=True)) ** 2).sum(axis=0)).max()
        assert 2 * r < 16384 * 0.004
        assert p.n_el == 2
