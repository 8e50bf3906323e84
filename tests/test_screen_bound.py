"""CPU check of the error bound behind the FP32 screen of the FP64 histogram
kernel: for pairs drawn from two atom boxes, the bin-unit distance computed
the kernel's way in FP32 (frame centred between the box centres, coordinates
rounded to float, fused sum of squares, square root off by up to one ulp)
stays within E(t) = rho*t + sigma of the exact distance, with rho = 11*2^-24
and sigma = 4*(r_I + r_J)/bin*2^-24 + 2^-16.  A plausible mistake in the bound
(dropping the box term, or the factor 2) fails on the far pairs."""
import numpy as np

RHO = 11.0 * 2.0 ** -24


def f32_fma(a, b, c):
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def screen_t(xi, xj, o, inv_bin, rng):
    Xi = ((xi - o) * inv_bin).astype(np.float32)
    Xj = ((xj - o) * inv_bin).astype(np.float32)
    d = (Xi - Xj).astype(np.float32)
    s = f32_fma(d[2], d[2], f32_fma(d[1], d[1], np.float32(d[0] * d[0])))
    t = np.sqrt(s, dtype=np.float32)
    # hardware square root: within one ulp of the rounded value
    return np.float32(t + rng.integers(-1, 2, size=t.shape) * np.spacing(t))


def test_screen_bound_holds_with_margin():
    rng = np.random.default_rng(3)
    worst = 0.0
    for trial in range(40):
        bin_ = [0.001, 0.005, 0.01, 0.05][trial % 4]
        half = rng.uniform(2.0, 40.0)                  # box half-edges
        ci = rng.uniform(-500, 500, 3)
        # overlapping boxes (short pairs, large frame coordinates) and far ones
        cj = ci + rng.normal(size=3) * (rng.uniform(0, 5) if trial % 2 else rng.uniform(0, 400))
        xi = ci[:, None] + rng.uniform(-half, half, (3, 4000))
        xj = cj[:, None] + rng.uniform(-half, half, (3, 4000))
        bi = np.concatenate([xi.min(axis=1), xi.max(axis=1)])
        bj = np.concatenate([xj.min(axis=1), xj.max(axis=1)])
        o = 0.25 * (bi[:3] + bi[3:] + bj[:3] + bj[3:])
        rI = 0.5 * np.linalg.norm(bi[3:] - bi[:3])
        rJ = 0.5 * np.linalg.norm(bj[3:] - bj[:3])
        inv = 1.0 / bin_
        t = screen_t(xi, xj, o[:, None], inv, rng).astype(np.float64)
        texact = np.sqrt(((xi - xj) ** 2).sum(axis=0)) / bin_
        sigma = 4.0 * (rI + rJ) * inv * 2.0 ** -24 + 2.0 ** -16
        E = RHO * t + sigma
        err = np.abs(t - texact)
        assert np.all(err < E), (trial, float((err / E).max()))
        # the bound without its safety factor of 2 still holds: the margin is real
        assert np.all(err < 0.5 * (RHO * t + 4.0 * (rI + rJ) * inv * 2.0 ** -24) + 2.0 ** -16)
        worst = max(worst, float((err / E).max()))
    assert worst > 0.05   # the bound is not vacuous by orders of magnitude
