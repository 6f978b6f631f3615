"""qmx_dot_noalloc (the overlap product with its chunk partials in a fixed
device array) against qmx_dot with a caller workspace, through the C ABI:
two calls in a row on the same GPU, the second with a smaller input, in both
precisions; the results must be bit-identical (same kernels, same partial
sums). A product whose partials exceed the fixed array is refused."""
import ctypes as ct

import numpy as np
import pytest
import torch

from qimpy_amd import _lib, rc

pytestmark = pytest.mark.gpu


def _operands(rng, batch, M, N, G, dtype):
    def crandn(shape):
        return torch.tensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape),
                            device="cuda").to(dtype)
    return crandn((batch, M, G)), crandn((batch, N, G))


def _dot_ws(code, A, B, out):
    b, M, G = A.shape
    N = B.shape[1]
    need = int(_lib.lib().qmx_dot_workspace_bytes(code, b, M, N, G))
    ws = torch.empty(max(need, 1), dtype=torch.uint8, device="cuda")
    _lib.call("qmx_dot", code, A.data_ptr(), G, M * G, B.data_ptr(), G, N * G, b, M, N, G,
              out.data_ptr(), N, M * N, 0, ws.data_ptr(), need, rc.stream_ptr())


def _dot_noalloc(code, A, B, out):
    b, M, G = A.shape
    N = B.shape[1]
    _lib.call("qmx_dot_noalloc", code, A.data_ptr(), G, M * G, B.data_ptr(), G, N * G, b, M, N,
              G, out.data_ptr(), N, M * N, 0, rc.stream_ptr())


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_noalloc_matches_workspace_dot(gpu, dtype):
    code = _lib.QMX_C64 if dtype == torch.complex64 else _lib.QMX_C128
    rng = np.random.default_rng(42)
    for shape in ((6, 144, 72, 7259), (3, 40, 24, 905)):  # the second call is smaller
        print(f"[noalloc] {dtype} batch {shape[0]}, {shape[1]} x {shape[2]}, G {shape[3]}", flush=True)
        A, B = _operands(rng, *shape, dtype)
        ref = torch.empty((shape[0], shape[1], shape[2]), dtype=torch.complex128, device="cuda")
        got = torch.full_like(ref, float("nan"))
        _dot_ws(code, A, B, ref)
        _dot_noalloc(code, A, B, got)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        # and against an FP64 product on the host:
        exact = np.einsum("bmg,bng->bmn", A.cpu().numpy().astype(np.complex128).conj(),
                          B.cpu().numpy().astype(np.complex128))
        tol = 1e-12 if dtype == torch.complex128 else 1e-6
        assert np.abs(got.cpu().numpy() - exact).max() / np.abs(exact).max() < tol


def test_noalloc_refuses_beyond_its_scratch(gpu):
    """A product whose chunk partials exceed the fixed array returns
    QMX_ERR_NOMEM with a message, and writes nothing."""
    cap = int(_lib.lib().qmx_dot_noalloc_max_bytes())
    assert cap == 32 << 20
    batch, M, N, G = 64, 256, 256, 60000
    need = int(_lib.lib().qmx_dot_workspace_bytes(_lib.QMX_C128, batch, M, N, G))
    assert need > cap
    out = torch.zeros((1,), dtype=torch.complex128, device="cuda")
    dummy = torch.zeros((1,), dtype=torch.complex128, device="cuda")
    status = _lib.lib().qmx_dot_noalloc(_lib.QMX_C128, ct.c_void_p(dummy.data_ptr()), G, M * G,
                                        ct.c_void_p(dummy.data_ptr()), G, N * G, batch, M, N, G,
                                        ct.c_void_p(out.data_ptr()), N, M * N, 0,
                                        ct.c_void_p(rc.stream_ptr()))
    assert status == 3  # QMX_ERR_NOMEM
    assert b"fixed scratch" in _lib.lib().qmx_last_error()
    torch.cuda.synchronize()
    assert not out.abs().any()
