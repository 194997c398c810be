"""The allocation-free evaluation entry point (minim_potential_eval_async_noalloc)
against the allocating one (minim_potential_eval_async), through the C ABI.

Both run the same kernels with the same reduction layout, so energies and
gradients must agree bit for bit.  Two calls in a row, the second on a
smaller field, check that the shared scratch array carries nothing from one
call into the next.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def field(grid, nf, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(0, 1, (int(np.prod(grid)), nf))
    return (u / u.sum(axis=1, keepdims=True)).reshape(-1)


def pf_handle(m, grid, nf, x, **kw):
    tension = [1.0, 0.8, 1.2] if nf == 3 else 1.0
    pf = m.PhaseField().setNFluid(nf).setGridSize(list(grid)).setSurfaceTension(tension).setInterfaceSize(2.0)
    if kw.get("volume_fixed"):
        pf.setVolumeFixed(True)
    return pf.handle(x)


def run(m, h, prec, x, n, noalloc):
    import torch

    A = m._capi
    L = A.load()
    dt = torch.float64 if prec == "fp64" else torch.float32
    xd = torch.from_numpy(x).to("cuda", dt)
    g = torch.full((n,), float("nan"), dtype=dt, device="cuda")
    e = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    fn = L.minim_potential_eval_async_noalloc if noalloc else L.minim_potential_eval_async
    rc = fn(h.ptr, 0 if prec == "fp64" else 1, C.c_void_p(xd.data_ptr()), n, None, C.c_double(0.0), 1,
            C.c_void_p(e.data_ptr()), C.c_void_p(g.data_ptr()), None)
    assert rc == 0, L.minim_last_error()
    torch.cuda.synchronize()
    return e.item(), g.cpu().numpy()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_equals_allocating_entry_point(gpu, prec):
    m = gpu
    cases = [((64, 48, 40), 3, 1), ((20, 12, 9), 3, 2)]  # the second call on a smaller field
    for grid, nf, seed in cases:
        print("case", prec, grid, nf, flush=True)
        x = field(grid, nf, seed)
        n = x.size
        h = pf_handle(m, grid, nf, x)
        e_ref, g_ref = run(m, h, prec, x, n, noalloc=False)  # also makes the tables resident
        e_na, g_na = run(m, h, prec, x, n, noalloc=True)
        assert e_na == e_ref, (e_na, e_ref)
        assert np.array_equal(g_na, g_ref)
        assert np.isfinite(g_na).all() and np.isfinite(e_na)


def test_noalloc_fixed_volume_and_binary(gpu):
    """The fixed-volume projection's reductions (up to 16 fluid sums) and a
    one-fluid field go through the same array."""
    m = gpu
    for grid, nf, kw, seed in [((32, 24, 16), 3, dict(volume_fixed=True), 3), ((50, 40, 1), 1, {}, 4)]:
        print("case", grid, nf, kw, flush=True)
        x = field(grid, nf, seed) if nf > 1 else np.random.default_rng(seed).uniform(-0.5, 0.5, int(np.prod(grid)))
        h = pf_handle(m, grid, nf, x, **kw)
        e_ref, g_ref = run(m, h, "fp64", x, x.size, noalloc=False)
        e_na, g_na = run(m, h, "fp64", x, x.size, noalloc=True)
        assert e_na == e_ref and np.array_equal(g_na, g_ref)


def test_noalloc_rejects_what_it_cannot_do(gpu):
    import torch

    m = gpu
    L = m._capi.load()
    grid, nf = (8, 8, 8), 3
    x = field(grid, nf, 5)
    h = pf_handle(m, grid, nf, x)
    xd = torch.from_numpy(x).cuda()
    g = torch.empty_like(xd)
    e = torch.empty(1, dtype=torch.float64, device="cuda")
    # tables not resident yet (no evaluation so far): a clear error, nothing launched
    rc = L.minim_potential_eval_async_noalloc(h.ptr, 0, C.c_void_p(xd.data_ptr()), x.size, None, C.c_double(0.0), 1,
                                              C.c_void_p(e.data_ptr()), C.c_void_p(g.data_ptr()), None)
    assert rc == 1 and b"resident" in L.minim_last_error()
    # another potential type
    lj = m.LjNd(3).handle(np.arange(12, dtype=float) * 1.3)
    rc = L.minim_potential_eval_async_noalloc(lj.ptr, 0, C.c_void_p(xd.data_ptr()), 12, None, C.c_double(0.0), 1,
                                              C.c_void_p(e.data_ptr()), None, None)
    assert rc == 1 and b"PhaseField only" in L.minim_last_error()
