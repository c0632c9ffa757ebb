"""GPU test of the entry point that allocates nothing
(i3g_stokes_residual_noalloc): through the C ABI on device buffers, against
i3g_stokes_residual on the same inputs, on two calls in a row (the second on
a smaller grid, so the shared scratch holds the first call's partial sums):
the residual equal to the bit, the norms to 1e-15 (the square roots are taken
on the device instead of the host)."""
import ctypes as C

import numpy as np
import pytest

import i3elvis_amd as I

pytestmark = pytest.mark.gpu


def _noalloc(S, x_planes, with_r=True):
    import torch
    lib = I.lib()
    lib.i3g_stokes_residual_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    N = S.N
    dx = torch.tensor(np.concatenate(x_planes), dtype=torch.float64, device="cuda")
    dr = torch.full((4 * N,), np.nan, dtype=torch.float64, device="cuda")
    dn = torch.zeros(2, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = lib.i3g_stokes_residual_noalloc(S.ctx, C.c_void_p(dx.data_ptr()),
                                         C.c_void_p(dr.data_ptr()) if with_r else None,
                                         C.c_void_p(dn.data_ptr()), None)
    torch.cuda.synchronize()
    I.check(rc, "stokes_residual_noalloc")
    r = dr.cpu().numpy()
    return [r[q * N:(q + 1) * N] for q in range(4)], dn.cpu().numpy()


def test_noalloc_residual_matches_residual(gpu):
    rng = np.random.default_rng(17)
    for n in (33, 17):  # second call smaller
        cfg = I.make_config(2, n, n, n, seed=3)
        print(f"case: residual_noalloc n={n}", flush=True)
        S = I.Stokes(cfg)
        S.set_coefficients()
        x = [rng.standard_normal(n ** 3) * s for s in (1e6, 1e-10, 1e-10, 1e-10)]
        r_ref, nrm_ref = S.residual(*x)
        r, nrm = _noalloc(S, x)
        for q in range(4):
            assert np.array_equal(r[q], r_ref[q]), (n, q)
        assert np.allclose(nrm, nrm_ref, rtol=1e-15, atol=0), (n, nrm, nrm_ref)
        _, nrm2 = _noalloc(S, x, with_r=False)
        assert np.allclose(nrm2, nrm_ref, rtol=1e-15, atol=0)
        S.close()
