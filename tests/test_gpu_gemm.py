"""qmx_gemm (rotations, back-projections, Gram-Schmidt updates) against an
FP64 product on the host, through the C ABI.

Shapes: those of the batched solver (2 nb -> nb rotations for nb = 32 and
72, projector back-projections P -> B), of the 64-atom solve (564 -> 282 over
16k plane waves) and of the 256-atom CheFSI (1024 -> 1024, a truncated basis
here), plus ragged edges (m, K and G not multiples of the tile sizes) and a
single output row. Tolerances (relative to the largest result entry):
complex128 on FP64 MFMA 1e-13; complex64 on FP32 MFMA 1e-5; BF16 x 3 split
precision 1e-4 (about 2^-16 per product, summed over K)."""
import ctypes as ct

import numpy as np
import pytest
import torch

from qimpy_amd import _lib, gemm, rc

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128
TOL = {"c128": 1e-13, "c64": 1e-5, "bf16x3": 1e-4}


def crandn(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


SHAPES = [  # (batch, Nin, Nout, G)
    (7, 144, 72, 7259), (5, 64, 32, 905), (1, 564, 282, 16135), (3, 72, 40, 3054),
    (2, 17, 1, 129), (1, 1024, 1024, 2000), (4, 256, 72, 1767), (1, 3, 5, 61),
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("kind", ["c128", "c64", "bf16x3"])
def test_rotate_matches_fp64(gpu, shape, kind):
    B, Nin, Nout, G = shape
    print(f"[rotate] {kind} batch {B}, {Nin} -> {Nout}, G {G}", flush=True)
    rng = np.random.default_rng(Nin * 7 + G)
    M = crandn(rng, (B, Nin, Nout)) / np.sqrt(Nin)
    W = crandn(rng, (B, Nin, G))
    O = crandn(rng, (B, Nout, G))
    dt = C128 if kind == "c128" else C64
    Mt = torch.tensor(M, device="cuda")  # complex128 matrix in every case
    Wt = torch.tensor(W, device="cuda").to(dt)
    Ot = torch.tensor(O, device="cuda").to(dt)
    Wr = Wt.cpu().numpy().astype(np.complex128)
    Or = Ot.cpu().numpy().astype(np.complex128)
    gemm.set_bf16x3(kind == "bf16x3")
    try:
        got = gemm.rotate(Mt, Wt)
        got2 = gemm.rotate(Mt, Wt, out=Ot.clone(), alpha=-0.5, beta=1.0)
    finally:
        gemm.set_bf16x3(False)
    ref = np.einsum("bkm,bkg->bmg", M, Wr)
    err = np.abs(got.cpu().numpy() - ref).max() / np.abs(ref).max()
    ref2 = Or - 0.5 * ref
    err2 = np.abs(got2.cpu().numpy() - ref2).max() / np.abs(ref2).max()
    print(f"  rel err {err:.2e}, accumulate {err2:.2e}", flush=True)
    assert err < TOL[kind] and err2 < TOL[kind]


@pytest.mark.parametrize("dtype", [C64, C128])
def test_rotate_strided_views(gpu, dtype):
    """Transposed and sliced matrices (M^T views, eigenvector column slices),
    operands and outputs that are row slices of a larger buffer, and a
    broadcast matrix over the batch."""
    rng = np.random.default_rng(3)
    K, nb, G = 4, 24, 777
    X = torch.tensor(crandn(rng, (K, 2 * nb, G)), device="cuda").to(dtype)
    V = torch.tensor(crandn(rng, (K, 2 * nb, 2 * nb)), device="cuda")
    Xr = X.cpu().numpy().astype(np.complex128)
    Vr = V.cpu().numpy()
    tol = 1e-13 if dtype == C128 else 1e-5
    # eigenvector column slice, output into the first half of a buffer:
    out = torch.zeros_like(X)
    gemm.rotate(V[:, :, :nb], X, out=out[:, :nb])
    ref = np.einsum("bkm,bkg->bmg", Vr[:, :, :nb], Xr)
    assert np.abs(out[:, :nb].cpu().numpy() - ref).max() / np.abs(ref).max() < tol
    assert not out[:, nb:].abs().any()
    # transposed view (op(M) = Ut with Ut stored row-major):
    Ut = V[:, :nb, :nb]
    got = gemm.rotate(Ut.transpose(-2, -1), X[:, nb:])
    ref = np.einsum("bmk,bkg->bmg", Vr[:, :nb, :nb], Xr[:, nb:])
    assert np.abs(got.cpu().numpy() - ref).max() / np.abs(ref).max() < tol
    # one matrix broadcast over the batch:
    M1 = V[0, :, :nb]
    got = gemm.rotate(M1, X)
    ref = np.einsum("km,bkg->bmg", Vr[0, :, :nb], Xr)
    assert np.abs(got.cpu().numpy() - ref).max() / np.abs(ref).max() < tol


def test_blas_entry_points(gpu):
    """qmx_gemm_{c64,c64_bf16x3,c128}: C = alpha op(A) B + beta C, row-major,
    batched, through ctypes."""
    rng = np.random.default_rng(9)
    batch, m, n, k = 3, 40, 1001, 56
    A = crandn(rng, (batch, k, m))  # transA: op(A) = A^T (m x k)
    Bm = crandn(rng, (batch, k, n))
    C0 = crandn(rng, (batch, m, n))
    lib = _lib.lib()
    for name, dt, tol in (("qmx_gemm_c128", C128, 1e-13), ("qmx_gemm_c64", C64, 1e-5),
                          ("qmx_gemm_c64_bf16x3", C64, 1e-4)):
        At, Bt = torch.tensor(A, device="cuda").to(dt), torch.tensor(Bm, device="cuda").to(dt)
        Ct = torch.tensor(C0, device="cuda").to(dt)
        ref = 2.0 * np.einsum("bkm,bkn->bmn", At.cpu().numpy().astype(np.complex128),
                              Bt.cpu().numpy().astype(np.complex128)) + 0.5 * Ct.cpu().numpy()
        _lib.call(name, 1, 0, m, n, k, ct.c_double(2.0), At.data_ptr(), m, k * m, Bt.data_ptr(),
                  n, k * n, ct.c_double(0.5), Ct.data_ptr(), n, m * n, batch, rc.stream_ptr())
        err = np.abs(Ct.cpu().numpy() - ref).max() / np.abs(ref).max()
        print(f"[{name}] rel err {err:.2e}", flush=True)
        assert err < tol
    # transposed B is refused with an error, not computed:
    with pytest.raises(RuntimeError, match="transposed B"):
        _lib.call("qmx_gemm_c64", 1, 1, m, n, k, ct.c_double(1.0), At.data_ptr(), m, k * m,
                  Bt.data_ptr(), n, k * n, ct.c_double(0.0), Ct.data_ptr(), n, m * n, batch,
                  rc.stream_ptr())


@pytest.mark.parametrize("n", [32, 72, 80, 128])
@pytest.mark.parametrize("kind", ["c128", "c64", "bf16x3"])
def test_rotate_in_place(gpu, n, kind):
    """out = C (in place, one block of output rows) equals the out-of-place
    rotation bit for bit, with and without accumulation; beyond the block
    limit the library refuses it."""
    dt = C128 if kind == "c128" else C64
    gemm.set_bf16x3(kind == "bf16x3")
    try:
        limit = gemm.inplace_rows_max(dt)
        print(f"[rotate in place] {kind} n {n} (limit {limit})", flush=True)
        rng = np.random.default_rng(n)
        M = torch.tensor(crandn(rng, (6, n, n)) / np.sqrt(n), device="cuda")
        W = torch.tensor(crandn(rng, (6, n, 1411)), device="cuda").to(dt)
        if n > limit:
            with pytest.raises(ValueError):
                gemm.rotate(M, W, out=W)
            return
        ref = gemm.rotate(M, W)
        ref2 = gemm.rotate(M, W, out=W.clone(), alpha=0.5, beta=-1.0)
        X = W.clone()
        gemm.rotate(M, X, out=X)
        X2 = W.clone()
        gemm.rotate(M, X2, out=X2, alpha=0.5, beta=-1.0)
        assert torch.equal(X, ref) and torch.equal(X2, ref2)
    finally:
        gemm.set_bf16x3(False)


@pytest.mark.parametrize("dtype", [C64, C128])
def test_rotate_past_resource_range(gpu, dtype):
    """A rotation whose long operand passes 2^30 bytes per batch takes the
    64-bit-addressed kernel; its columns equal those of the same rotation
    over a compact slice (the buffer-resource kernel, same arithmetic
    order), bit for bit, and a transposed M takes the same route."""
    K, m = 16, 40
    G = (1 << 30) // (K * torch.empty((), dtype=dtype).element_size()) + 1000
    print(f"[rotate_range] {dtype} K {K} m {m} G {G}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    W = torch.randn((1, K, G), dtype=dtype, device="cuda", generator=g)
    M = torch.randn((1, m, K), dtype=C128, device="cuda", generator=g)
    out = gemm.rotate(M.transpose(-2, -1), W)
    for a, b in ((0, 777), (G - 700, G)):
        ref = gemm.rotate(M.transpose(-2, -1), W[:, :, a:b].contiguous())
        assert torch.equal(out[:, :, a:b], ref)
    outT = gemm.rotate(M.transpose(-2, -1).contiguous(), W)
    assert torch.equal(outT[:, :, :500],
                       gemm.rotate(M.transpose(-2, -1).contiguous(), W[:, :, :500].contiguous()))
