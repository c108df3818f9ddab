"""mpsa_dot_noalloc -- the dot product of mpsa_dot for callers that cannot
allocate (its partial sums live in a file-scope device array, its result goes
to a caller device pointer) -- against mpsa_dot through the C ABI: bitwise the
same (re, im) on two calls in a row, the second with a smaller vector, real
and complex."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cplx", [0, 1])
def test_dot_noalloc_equals_dot(gpu, cplx):
    from mpstools_amd import native
    torch = gpu
    L = native.lib()
    L.mpsa_dot_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p]
    for n in (1_000_003, 777):
        print("[case] dot_noalloc", n, cplx, flush=True)
        x = torch.randn(n * (2 if cplx else 1), dtype=torch.float64, device="cuda")
        y = torch.randn(n * (2 if cplx else 1), dtype=torch.float64, device="cuda")
        out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
        rc = L.mpsa_dot_noalloc(C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), n, cplx,
                                C.c_void_p(out.data_ptr()), native._stream())
        assert rc == 0, L.mpsa_last_error().decode()
        torch.cuda.synchronize()
        ref = np.zeros(2)
        native._check(L.mpsa_dot(C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()), n, cplx, native._ptr(ref),
                                 native._stream()))
        assert np.array_equal(out.cpu().numpy(), ref)
        xs, ys = x.cpu().numpy(), y.cpu().numpy()
        if cplx:
            want = np.vdot(ys[0::2] + 1j * ys[1::2], xs[0::2] + 1j * xs[1::2])
            assert abs(complex(*ref) - want) <= 1e-9 * n ** 0.5
        else:
            assert abs(ref[0] - xs @ ys) <= 1e-9 * n ** 0.5
