"""The allocation-free Debye entry point (xs_debye_pairsum_noalloc) against
xs_debye_pairsum on the same device inputs: two calls in a row, the second
with a smaller structure and mesh, so that the shared scratch keeps the first
call's tables and must not leak into the second.  FP64 pair sums: the two
entry points run the same kernels with the same tables, so they agree to
rounding (bar 1e-12 of max |S|; the atomics' order differs)."""
import numpy as np
import pytest
try:
    import torch  # noqa: F401  (first import can take a minute on a fresh box)
except ImportError:
    torch = None

import xansons_hip as xh
from xansons_hip import structures as st

pytestmark = pytest.mark.gpu


def run_pair(ctx, ens, q):
    dev = torch.device("cuda", 0)
    xyz = torch.from_numpy(np.ascontiguousarray(ens.xyz64())).to(dev)
    qd = torch.from_numpy(np.ascontiguousarray(q)).to(dev)
    P = ens.n_el * (ens.n_el + 1) // 2
    Sa = torch.full((P * q.size,), 7.0, dtype=torch.float64, device=dev)  # stale values
    Sb = torch.empty(P * q.size, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    print(f"noalloc: {ens.n} atoms, {ens.n_el} elements, {q.size} q", flush=True)
    ctx.debye_pairsum_noalloc_dev(xyz.data_ptr(), ens.n, ens.el_count, qd.data_ptr(), q.size,
                                  Sa.data_ptr())
    ctx.synchronize()
    ctx.debye_pairsum_dev(xyz.data_ptr(), ens.n, ens.el_count, qd.data_ptr(), q.size, Sb.data_ptr())
    ctx.synchronize()
    a = Sa.cpu().numpy().reshape(P, q.size)
    b = Sb.cpu().numpy().reshape(P, q.size)
    return a, b


def check(a, b):
    scale = np.abs(b).max(axis=1, keepdims=True)
    err = float((np.abs(a - b) / np.maximum(scale, 1e-300)).max())
    print(f"  max |dS| / max |S| = {err:.3e}", flush=True)
    assert err <= 1e-12


def test_noalloc_two_calls_second_smaller():
    with xh.Context(0) as ctx:
        big = st.disordered_cubic(n=14, natoms=None)      # 2,744 atoms, 3 elements
        check(*run_pair(ctx, big, st.q_mesh(0.5, 10.0, 1500)))
        small = st.fcc_sphere(radius=8.0, seed=42)        # 135 atoms, 1 element
        check(*run_pair(ctx, small, st.q_mesh(0.3, 8.0, 200)))


def test_noalloc_non_uniform_mesh_and_single_atom_types():
    with xh.Context(0) as ctx:
        ens = st.random_cluster([1, 40, 33], box=9.0, min_dist=0.9, seed=43)
        q = np.sort(np.random.default_rng(3).uniform(0.2, 12.0, 300))
        check(*run_pair(ctx, ens, q))
        check(*run_pair(ctx, ens, st.q_mesh(0.5, 10.0, 1024)))


def test_noalloc_rejects_beyond_scratch():
    """Tables beyond the 32 MiB scratch: a clear XS_ERR_RANGE from the host
    check, before anything is launched (the pointers are never read)."""
    with xh.Context(0) as ctx:
        with pytest.raises(xh.XSError) as e:
            ctx.debye_pairsum_noalloc_dev(0x1000, 70000, [70000], 0x1000, 16, 0x1000)
        assert e.value.code == -5
        assert "scratch" in str(e.value)
