"""Every launch choice the autotuner can set gives the right result: the
overlap product (qmx_dot) for every candidate blocking and chunk count, the
rotation (qmx_gemm) for every output-block height, and the plane pass for
each kernel choice, against FP64 references (and the oracle for the plane
pass), to the tolerances of the untuned kernels. Also: the shipped cache
loads for this device and a cache of another device is ignored."""
import json

import numpy as np
import pytest
import torch

from oracle.basis import Basis as OBasis
from oracle.hamiltonian import apply_potential as o_apply_potential
from qimpy_amd import autotune, gemm, systems
from qimpy_amd.basis import Basis
from qimpy_amd.wavefunction import Wavefunction

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128


@pytest.fixture(autouse=True)
def _clean_table():
    autotune.clear()
    yield
    autotune.clear()
    autotune._loaded_default = False


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("shape", [(16, 64, 32, 3000), (5, 144, 72, 2100), (3, 33, 17, 700)])
def test_dot_every_candidate(gpu, dtype, shape):
    batch, M, N, G = shape
    g = torch.Generator().manual_seed(M + N)
    A = torch.randn(batch, M, G, dtype=C128, generator=g)
    B = torch.randn(batch, N, G, dtype=C128, generator=g)
    ref = torch.matmul(A.conj(), B.transpose(-2, -1))
    scale = float(A.abs().max() * B.abs().max()) * G
    tol = 1e-12 if dtype == C128 else 2e-6
    Ad, Bd = A.to(dtype).cuda(), B.to(dtype).cuda()
    key = autotune.dot_key(0 if dtype == C64 else 1, batch, M, N, G)
    cands = autotune.dot_candidates(batch, M, N, G)
    print(f"[autotune dot] {shape} {dtype}: {len(cands)} candidates", flush=True)
    for cand in cands:
        autotune._set(autotune.TUNE_DOT, key, cand)
        got = gemm.dot128(Ad, Bd).cpu()
        err = float((got - ref).abs().max()) / scale
        assert err < tol, (cand, err)


@pytest.mark.parametrize("dtype", [C64, C128])
def test_rotate_every_block_height(gpu, dtype):
    batch, K, m, G = 4, 144, 72, 3000
    g = torch.Generator().manual_seed(3)
    Mx = torch.randn(batch, K, m, dtype=C128, generator=g)
    C = torch.randn(batch, K, G, dtype=C128, generator=g)
    ref = torch.matmul(Mx.transpose(-2, -1), C)
    scale = float(Mx.abs().max() * C.abs().max()) * K
    tol = 1e-12 if dtype == C128 else 2e-6
    Md, Cd = Mx.to(dtype).cuda(), C.to(dtype).cuda()
    code = 0 if dtype == C64 else 1
    key = autotune.rotate_key(code, code, 0, m, K, G)
    for mt in range(1, (5 if dtype == C128 else 8) + 1):
        autotune._set(autotune.TUNE_ROTATE, key, (mt,))
        got = gemm.rotate(Md, Cd).cpu()
        err = float((got - ref).abs().max()) / scale
        print(f"[autotune rotate] {dtype} MT={mt}: {err:.1e}", flush=True)
        assert err < tol


@pytest.mark.parametrize("dtype", [C64, C128])
def test_plane_every_kernel(gpu, dtype):
    spec = systems.small_cell(shape=(30, 30, 30), n_bands=4, kmesh=(1, 1, 2), triclinic=False,
                              seed=130)
    basis = Basis.from_spec(spec)
    ob = OBasis(spec)
    rng = np.random.default_rng(2)
    C = rng.standard_normal((len(ob.k), 3, ob.n_tot)) + 1j * rng.standard_normal(
        (len(ob.k), 3, ob.n_tot))
    ob.constrain(C)
    ref = o_apply_potential(ob, spec.V, C)
    wf = Wavefunction(basis, coeff=torch.tensor(C, device="cuda")[None, :, :, None, :].to(
        dtype).contiguous())
    V = torch.as_tensor(spec.V, device="cuda")
    key = (0 if dtype == C64 else 1, 30, 30, 30)
    for choice in (0, 1, 2):
        autotune._set(autotune.TUNE_PLANE, key, (choice,))
        got = basis.apply_potential(V, wf).coeff[0, :, :, 0].cpu().numpy()
        err = np.abs(got - ref).max() / np.abs(ref).max()
        print(f"[autotune plane] {dtype} choice {choice}: {err:.1e}", flush=True)
        assert err < (1e-12 if dtype == C128 else 2e-5)


def test_cache_device_match(gpu, tmp_path):
    assert autotune.load_default() > 0  # the shipped cache of this device model
    autotune.clear()
    data = json.loads(autotune.default_path().read_text())
    data["device"] = "other_device"
    p = tmp_path / "other.json"
    p.write_text(json.dumps(data))
    assert autotune.load(p) == 0
    assert int(autotune._lib.lib().qmx_tune_count()) == 0
