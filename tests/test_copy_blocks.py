"""The restack's batched row-block copy (qmx_copy_blocks): descriptors built
by batched._BlockCopies, checked on the CPU by applying them to host
memory the way the kernel does, and on the GPU against torch's copies."""
import ctypes

import numpy as np
import pytest
import torch

from qimpy_amd import batched

C64, C128, F64 = torch.complex64, torch.complex128, torch.float64


def cases(device):
    g = torch.Generator().manual_seed(3)

    def rnd(shape, dt):
        x = torch.randn(*shape, 2 if dt.is_complex else 1, generator=g, dtype=torch.float64)
        x = torch.view_as_complex(x) if dt.is_complex else x[..., 0]
        return x.to(dt).to(device)

    out = []
    # kept wavefunction rows into a longer basis (zero past the old length)
    X_old = rnd((5, 12, 37), C64)
    X_new = rnd((7, 12, 48), C64)
    out.append((X_new[1:4, :6], X_old[2:5, :6]))
    Y_old = rnd((4, 6, 30), C128)
    Y_new = rnd((4, 6, 33), C128)
    out.append((Y_new[1:3], Y_old[0:2]))
    # eigenvalue rows, a broadcast coupling row, a potential, zero fill
    e_old, e_new = rnd((9, 5), F64), rnd((9, 5), F64)
    out.append((e_new[3:7], e_old[1:5]))
    D = rnd((6, 8), F64)
    out.append((D[2:5], rnd((5,), F64)))
    V = rnd((3, 1000), F64)
    out.append((V[1], rnd((1000,), F64)))
    out.append((D[0:2], None))
    # complex128 projectors into a complex64 bucket, then its zero rows
    beta = rnd((4, 5, 40), C64)
    out.append((beta[1:3, :3], rnd((2, 3, 31), C128)))
    out.append((beta[1:3, 3:], None))
    return out


def expected(pairs):
    res = []
    for dst, src in pairs:
        e = torch.zeros_like(dst)
        if src is not None:
            s = src.expand(*dst.shape[:-1], src.shape[-1])
            e[..., : s.shape[-1]] = s.to(dst.dtype)
        res.append(e)
    return res


def apply_on_host(rows):
    """Each descriptor as k_copy_blocks applies it, on host memory."""
    for src, dst, n0, n1, sc, dc, ss0, ss1, ds0, ds1, rpw, fw, mode, _ in rows:
        st, dt = (ctypes.c_uint64, ctypes.c_uint64) if mode == 0 else (ctypes.c_double, ctypes.c_float)
        for i0 in range(n0):
            for i1 in range(n1):
                o = np.ctypeslib.as_array((dt * dc).from_address(
                    dst + (i0 * ds0 + i1 * ds1) * ctypes.sizeof(dt)))
                o[:] = 0
                if src:
                    s = np.ctypeslib.as_array((st * sc).from_address(
                        src + (i0 * ss0 + i1 * ss1) * ctypes.sizeof(st)))
                    o[:sc] = s


def test_copy_descriptors_on_host():
    pairs = cases("cpu")
    want = expected(pairs)
    base_rows = pairs[0][0]._base[4:].clone()
    cp = batched._BlockCopies()
    for dst, src in pairs:
        cp.add(dst, src)
    rows = cp.rows
    a = np.array(rows, dtype=np.int64)
    wg = -(-(a[:, 2] * a[:, 3]) // a[:, 10])
    assert np.all(wg >= 1)
    apply_on_host(rows)
    for (dst, _), w in zip(pairs, want):
        assert torch.equal(dst, w)
    # rows of the bases outside the blocks are untouched
    assert torch.equal(pairs[0][0]._base[4:], base_rows)


def test_copy_rejects():
    cp = batched._BlockCopies()
    with pytest.raises(ValueError):
        cp.add(torch.zeros(4, 3), torch.zeros(4, 5))
    with pytest.raises(ValueError):
        cp.add(torch.zeros(4, 6)[:, ::2])


@pytest.mark.gpu
def test_copy_blocks_gpu(gpu):
    """The kernel against torch's own copies, bit for bit; padding rows
    beside the blocks keep their values."""
    torch.manual_seed(0)
    pairs = cases("cuda")
    want = expected(pairs)
    base = pairs[0][0]._base
    keep_rows = base[4:].clone()
    cp = batched._BlockCopies()
    for dst, src in pairs:
        cp.add(dst, src)
    cp.run()
    torch.cuda.synchronize()
    for (dst, _), w in zip(pairs, want):
        assert torch.equal(dst, w)
    assert torch.equal(base[4:], keep_rows)
