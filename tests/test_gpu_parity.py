"""GPU parity of the HIP path against the FP64 oracle (MI355X).

Tolerances: complex128 kernels match the oracle to ~1e-12 relative; complex64
kernels to ~1e-5 relative (a few float32 ulps through two 3D transforms);
solver eigenvalues within 1e-8 Ha of the oracle's FP64 path for the plain
form and 1e-6 Ha for the mixed form.
"""
import numpy as np
import pytest
import torch

from oracle.basis import Basis as OBasis
from oracle.davidson import davidson as o_davidson
from oracle.davidson import initial_state as o_initial
from oracle.dense import dense_eigenvalues
from oracle.hamiltonian import Hamiltonian as OHam
from oracle.hamiltonian import apply_potential as o_apply_potential
from oracle.init import randomize as o_randomize
from qimpy_amd import systems
from qimpy_amd.electrons import electrons_from_spec
from qimpy_amd.grid import Grid
from qimpy_amd.linalg import eighg
from qimpy_amd.wavefunction import Wavefunction

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else b
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def rand_coeff(ob, n_bands, seed=0):
    rng = np.random.default_rng(seed)
    C = rng.standard_normal((len(ob.k), n_bands, ob.n_tot)) + 1j * rng.standard_normal(
        (len(ob.k), n_bands, ob.n_tot))
    ob.constrain(C)
    return C


def to_wf(basis, C, dtype):
    return Wavefunction(basis, coeff=torch.tensor(C, device="cuda")[None, :, :, None, :].to(dtype).contiguous())


@pytest.mark.parametrize("shape", [(36, 36, 36), (24, 30, 40), (6, 10, 15), (96, 96, 96), (144, 20, 60)])
@pytest.mark.parametrize("dtype", [C128, C64])
def test_fft3d(gpu, shape, dtype):
    print(f"[fft3d] {shape} {dtype}", flush=True)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2,) + shape) + 1j * rng.standard_normal((2,) + shape)
    g = Grid(shape)
    xt = torch.tensor(x, device="cuda").to(dtype)
    tol = 1e-13 if dtype == C128 else 3e-6
    assert rel(g.fft(xt), np.fft.fftn(x, axes=(1, 2, 3)) / np.prod(shape)) < tol
    assert rel(g.ifft(xt), np.fft.ifftn(x, axes=(1, 2, 3)) * np.prod(shape)) < tol
    # round trip (QimPy normalisation: ifft(fft(x)) = x)
    assert rel(g.ifft(g.fft(xt)), x) < tol


def test_fft_real_dispatch(gpu):
    shape = (12, 10, 8)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape)
    g = Grid(shape)
    xh = g.fft(torch.tensor(x, device="cuda"))
    assert tuple(xh.shape) == g.shapeH
    assert rel(xh, np.fft.rfftn(x) / x.size) < 1e-13
    assert rel(g.ifft(xh), x) < 1e-13


@pytest.fixture(scope="module")
def small_el(gpu):
    spec = systems.small_cell(n_bands=5)
    return spec, OBasis(spec), electrons_from_spec(spec)


@pytest.fixture(scope="module")
def cfg0_el(gpu):
    spec = systems.make_cell(0)
    return spec, OBasis(spec), electrons_from_spec(spec)


def test_basis_matches_oracle(small_el, cfg0_el):
    for spec, ob, el in (small_el, cfg0_el):
        b = el.basis
        assert np.array_equal(b.n_np, ob.n)
        assert np.array_equal(b.fft_index_np, ob.fft_index)
        assert rel(b.get_ke(), ob.KE) < 1e-14


@pytest.mark.parametrize("which", ["small", "cfg0"])
@pytest.mark.parametrize("dtype", [C128, C64])
def test_apply_potential(which, dtype, small_el, cfg0_el):
    spec, ob, el = small_el if which == "small" else cfg0_el
    print(f"[apply_potential] {which} {dtype}", flush=True)
    C = rand_coeff(ob, 7, seed=3)
    ref = o_apply_potential(ob, spec.V, C)
    got = el.basis.apply_potential(el.V_ks, to_wf(el.basis, C, dtype)).coeff[0, :, :, 0]
    assert rel(got, ref) < (1e-12 if dtype == C128 else 2e-5)
    kin = el.basis.apply_ke_potential(el.V_ks, to_wf(el.basis, C, dtype)).coeff[0, :, :, 0]
    assert rel(kin, ref + C * ob.KE[:, None, :]) < (1e-12 if dtype == C128 else 2e-5)


def _cell_with_box(g):
    i = 0
    while True:
        spec = systems.make_cfg2_cell(i, cfg=3)
        if int(spec.shape[0]) == g and len(np.asarray(spec.k).reshape(-1, 3)) <= 4:
            return spec
        i += 1


@pytest.mark.parametrize("g", [24, 30, 40, 48, 60])
@pytest.mark.parametrize("dtype", [C128, C64])
def test_apply_potential_boxes(gpu, g, dtype):
    """The fused plane pass on every box size of the batched workloads (the
    composite-radix transforms 12x2, 15x2, 10x4, 16x3, 15x4 and their plain
    fallbacks) against the oracle's convolution by FFT."""
    spec = _cell_with_box(g)
    print(f"[apply_potential_boxes] {g}^3 {dtype}", flush=True)
    ob = OBasis(spec)
    el = electrons_from_spec(spec)
    C = rand_coeff(ob, 3, seed=g)
    ref = o_apply_potential(ob, spec.V, C)
    got = el.basis.apply_potential(el.V_ks, to_wf(el.basis, C, dtype)).coeff[0, :, :, 0]
    assert rel(got, ref) < (1e-12 if dtype == C128 else 2e-5)


@pytest.mark.parametrize("g", [24, 36, 48, 60])
@pytest.mark.parametrize("dtype", [C128, C64])
def test_collect_density_boxes(gpu, g, dtype):
    """Basis.collect_density (qmx_collect_density) against the oracle on the
    box sizes of the batched workloads: fractional and zero fillings,
    complex128 to 1e-12 and complex64 to 1e-5 of the peak density; the cell
    integral equals sum f w (normalised bands)."""
    from oracle.density import collect_density as o_density
    spec = _cell_with_box(g)
    print(f"[collect_density] {g}^3 {dtype}", flush=True)
    ob = OBasis(spec)
    el = electrons_from_spec(spec)
    C = rand_coeff(ob, 7, seed=g + 1)
    C /= np.linalg.norm(C, axis=-1, keepdims=True)
    rng = np.random.default_rng(g)
    f = rng.uniform(0.0, 1.0, (len(ob.k), 7))
    f[:, 3] = 0.0  # an empty band
    ref = o_density(ob, C, f)
    got = el.basis.collect_density(to_wf(el.basis, C, dtype), torch.tensor(f)[None])
    assert got.shape == (1,) + ob.shape and got.dtype == torch.float64
    assert rel(got[0], ref) < (1e-12 if dtype == C128 else 1e-5)
    charge = float(got[0].sum()) * ob.volume / ob.N
    assert abs(charge - (f * ob.w_sk[:, None]).sum()) < (1e-10 if dtype == C128 else 1e-5)


def test_collect_density_spins_blocks_empty(cfg0_el):
    """Two spin channels give [n_tot, M_z]; a tiny workspace (many band
    blocks) gives the same density; all-zero fillings give zero."""
    from oracle.density import collect_density as o_density
    import qimpy_amd.basis as qb
    spec, ob, el = cfg0_el
    b = el.basis
    C0, C1 = rand_coeff(ob, 6, seed=21), rand_coeff(ob, 6, seed=22)
    rng = np.random.default_rng(7)
    f = rng.uniform(0.0, 1.0, (2, len(ob.k), 6))
    coeff = torch.stack([to_wf(b, C0, C128).coeff[0], to_wf(b, C1, C128).coeff[0]])
    w_sk = ob.w_sk.copy()
    try:
        ob.w_sk = w_sk / 2  # spin weight 1 with two channels (_basis.py:137)
        b.w_sk = b.w_sk / 2
        r0, r1 = o_density(ob, C0, f[0]), o_density(ob, C1, f[1])
        got = b.collect_density(coeff, torch.tensor(f))
        assert rel(got[0], r0 + r1) < 1e-12 and rel(got[1], r0 - r1) < 1e-12
        cap = qb.WORKSPACE_CAP
        try:
            qb.WORKSPACE_CAP = 1
            blocked = b.collect_density(coeff, torch.tensor(f))
        finally:
            qb.WORKSPACE_CAP = cap
        assert rel(blocked, got) < 1e-13
        zero = b.collect_density(coeff, torch.zeros(2, len(ob.k), 6, dtype=torch.float64))
        assert float(zero.abs().max()) == 0.0
    finally:
        ob.w_sk = w_sk
        b.w_sk = b.w_sk * 2


def _spinor_density_check(b, ob, dtype, n_bands, seed):
    """need_Mvec = True and False on random spinor bands against the oracle:
    n = rho_0 + rho_1, M_x = 2 Re X, M_y = 2 Im X, M_z = rho_0 - rho_1."""
    from oracle.density import collect_density as o_density
    from oracle.density import collect_density_cross as o_cross
    C0, C1 = rand_coeff(ob, n_bands, seed=seed), rand_coeff(ob, n_bands, seed=seed + 1)
    rng = np.random.default_rng(seed)
    f = rng.uniform(0.0, 1.0, (len(ob.k), n_bands))
    f[:, 1] = 0.0  # an empty band
    coeff = torch.cat([to_wf(b, C0, dtype).coeff, to_wf(b, C1, dtype).coeff], dim=3)
    r0, r1 = o_density(ob, C0, f), o_density(ob, C1, f)
    X = o_cross(ob, C0, C1, f)
    tol = 1e-12 if dtype == C128 else 1e-5
    got = b.collect_density(coeff, torch.tensor(f)[None], need_Mvec=True)
    assert got.shape == (4,) + ob.shape and got.dtype == torch.float64
    peak = np.abs(r0 + r1).max()
    for i, want in enumerate((r0 + r1, 2 * X.real, 2 * X.imag, r0 - r1)):
        err = np.abs(got[i].cpu().numpy() - want).max() / peak
        print(f"[spinor density] component {i}: {err:.2e}", flush=True)
        assert err < tol
    plain = b.collect_density(coeff, torch.tensor(f)[None])
    assert plain.shape == (1,) + ob.shape
    assert np.abs(plain[0].cpu().numpy() - (r0 + r1)).max() / peak < tol
    # one band with spin along +x, (phi, phi) / sqrt 2, then +y, (phi, i phi) / sqrt 2
    phi = C0[:, :1] / np.sqrt(2)
    fx = np.zeros((len(ob.k), 1))
    fx[0, 0] = 1.0
    for spin, axis in ((1.0, 1), (1j, 2)):
        cs = torch.cat([to_wf(b, phi, dtype).coeff, to_wf(b, spin * phi, dtype).coeff], dim=3)
        m = b.collect_density(cs, torch.tensor(fx)[None], need_Mvec=True).cpu().numpy()
        for i in (1, 2, 3):
            want = m[0] if i == axis else 0.0
            assert np.abs(m[i] - want).max() < tol * np.abs(m[0]).max()


@pytest.mark.parametrize("g", [24, 48, 60])
@pytest.mark.parametrize("dtype", [C128, C64])
def test_collect_density_spinor(gpu, g, dtype):
    """Spinor density with and without need_Mvec (qmx_collect_density per
    component, qmx_collect_density_cross for M_x, M_y) against the oracle,
    complex128 to 1e-12 and complex64 to 1e-5 of the peak density; a band
    polarised along +x or +y gives |M| = n along that axis only."""
    spec = _cell_with_box(g)
    print(f"[collect_density spinor] {g}^3 {dtype}", flush=True)
    ob = OBasis(spec)
    el = electrons_from_spec(spec)
    _spinor_density_check(el.basis, ob, dtype, 5, seed=g + 3)


def test_collect_density_spinor_large_plane(gpu):
    """96^3 box (512-thread plane kernels): the complex64 cross term on the
    plane pass and the complex128 one (two 96 x 97 planes exceed LDS) on the
    separate passes, against the oracle."""
    spec = systems.make_cell(1)
    print("[collect_density spinor] 96^3", flush=True)
    ob = OBasis(spec)
    el = electrons_from_spec(spec)
    b = el.basis
    _spinor_density_check(b, ob, C64, 3, seed=11)
    _spinor_density_check(b, ob, C128, 3, seed=12)


def test_apply_local_blocking_inplace_accumulate(cfg0_el):
    """Tiny workspace (many band blocks), out == C, and accumulate."""
    spec, ob, el = cfg0_el
    b = el.basis
    C = rand_coeff(ob, 9, seed=4)
    ref = o_apply_potential(ob, spec.V, C)
    c = to_wf(b, C, C128).coeff
    import qimpy_amd.basis as qb

    cap = qb.WORKSPACE_CAP
    try:
        qb.WORKSPACE_CAP = 1
        b._workspace = None
        out = b.apply_local(c, el.V_ks, with_ke=False)
    finally:
        qb.WORKSPACE_CAP = cap
        b._workspace = None
    assert rel(out[0, :, :, 0], ref) < 1e-12
    acc = c.clone()
    b.apply_local(c, el.V_ks, with_ke=False, out=acc, accumulate=True)
    assert rel(acc[0, :, :, 0], ref + C) < 1e-12
    inplace = c.clone()
    b.apply_local(inplace, el.V_ks, with_ke=True, out=inplace)
    assert rel(inplace[0, :, :, 0], ref + C * ob.KE[:, None, :]) < 1e-12


def test_randomize_matches_oracle(small_el, cfg0_el):
    for spec, ob, el in (small_el, cfg0_el):
        for (b0, b1) in ((0, 6), (2, 5)):
            C = np.zeros((len(ob.k), 6, ob.n_tot), complex)
            o_randomize(ob, C, seed=11, b_start=b0, b_stop=b1)
            W = Wavefunction(el.basis, n_bands=6)
            W.randomize(seed=11, b_start=b0, b_stop=b1)
            assert np.abs(W.coeff[0, :, :, 0].cpu().numpy() - C).max() < 1e-13


def test_band_reductions_and_precond(cfg0_el):
    spec, ob, el = cfg0_el
    C = rand_coeff(ob, 5, seed=5)
    HC = rand_coeff(ob, 5, seed=6)
    W = to_wf(el.basis, C, C128)
    norm = W.band_norm()[0].cpu().numpy()
    assert rel(norm, np.sqrt((np.abs(C) ** 2).sum(-1))) < 1e-14
    ke = W.band_ke()[0].cpu().numpy()
    assert rel(ke, np.einsum("kbg,kg->kb", np.abs(C) ** 2, ob.KE)) < 1e-14
    eig = np.random.default_rng(7).standard_normal((len(ob.k), 5))
    d = el.diagonalize
    # The fused residual + preconditioner pass (KEref = band_ke of C) and the
    # Python preconditioner against the oracle's restatement, whose
    # conventions the reference's own vectors pin (test_golden_reference).
    from oracle.davidson import precondition as o_precondition
    Cexp, nrm = d._expansion(W, to_wf(el.basis, HC, C128), torch.tensor(eig, device="cuda")[None])
    ref = o_precondition(ob, HC - C * eig[..., None], ke)
    assert rel(Cexp.coeff[0, :, :, 0], ref) < 1e-14
    assert rel(nrm[0], np.sqrt((np.abs(ref) ** 2).sum(-1))) < 1e-14
    rn = d.residual_norm[0].cpu().numpy()
    assert rel(rn, np.sqrt((np.abs(HC - C * eig[..., None]) ** 2).sum(-1))) < 1e-14
    P = d.precondition(to_wf(el.basis, HC, C128), torch.tensor(ke, device="cuda")[None])
    assert rel(P.coeff[0, :, :, 0], o_precondition(ob, HC, ke)) < 1e-14


@pytest.mark.parametrize("dtype", [C128, C64])
def test_precond_residual_normalize(cfg0_el, dtype):
    """qmx_precond_residual_normalize (the batched step's residual pass):
    the preconditioned residual of the oracle, scaled to unit norm per band
    and zero past each k's basis, with the norms before scaling. Complex64
    evaluates the preconditioner factor in FP32 (tolerance 1e-5)."""
    from oracle.davidson import precondition as o_precondition
    from qimpy_amd import _lib, rc
    from qimpy_amd.basis import DTYPE_CODE

    spec, ob, el = cfg0_el
    K, B, G = len(ob.k), 5, ob.n_tot
    C = rand_coeff(ob, B, seed=8)
    HC = rand_coeff(ob, B, seed=9)
    eig = np.random.default_rng(10).standard_normal((K, B))
    KEref = np.einsum("kbg,kg->kb", np.abs(C) ** 2, ob.KE)
    ref = o_precondition(ob, HC - C * eig[..., None], KEref)
    nref = np.sqrt((np.abs(ref) ** 2).sum(-1))
    dev = "cuda"
    Ct = torch.tensor(C, device=dev).to(dtype).contiguous()
    HCt = torch.tensor(HC, device=dev).to(dtype).contiguous()
    out = torch.full_like(Ct, 7.0)  # padding must come back zero
    norm2 = torch.empty((K, B), dtype=torch.float64, device=dev)
    rnorm2 = torch.empty_like(norm2)
    eig_t = torch.tensor(eig, device=dev)
    KEref_t = torch.tensor(KEref, device=dev)
    ke_t = torch.tensor(ob.KE, device=dev).contiguous()
    n32 = torch.tensor(ob.n, dtype=torch.int32, device=dev)
    _lib.call("qmx_precond_residual_normalize", DTYPE_CODE[dtype], Ct.data_ptr(), HCt.data_ptr(),
              G, B * G, eig_t.data_ptr(), KEref_t.data_ptr(), ke_t.data_ptr(), G,
              out.data_ptr(), G, B * G, K, B, G, norm2.data_ptr(), rnorm2.data_ptr(),
              n32.data_ptr(), rc.stream_ptr())
    tol = 1e-13 if dtype == C128 else 1e-5
    got = out.cpu().numpy()
    assert rel(got, ref / nref[..., None]) < tol
    assert rel(norm2.sqrt().cpu().numpy(), nref) < tol
    rn = np.sqrt((np.abs(HC - C * eig[..., None]) ** 2).sum(-1))
    assert rel(rnorm2.sqrt().cpu().numpy(), rn) < (1e-14 if dtype == C128 else 1e-6)
    assert not np.any(got[ob.pad_mask[:, None, :].repeat(B, axis=1)])


@pytest.mark.parametrize("dtype", [C128, C64])
def test_precond_residual(cfg0_el, dtype):
    """qmx_precond_residual (the residual pass of both solvers): the oracle's
    preconditioned residual with its norm and the unpreconditioned residual
    norm. Complex64 evaluates the preconditioner factor in FP32 (tolerance
    1e-5)."""
    from oracle.davidson import precondition as o_precondition
    from qimpy_amd import _lib, rc
    from qimpy_amd.basis import DTYPE_CODE

    spec, ob, el = cfg0_el
    K, B, G = len(ob.k), 5, ob.n_tot
    C = rand_coeff(ob, B, seed=18)
    HC = rand_coeff(ob, B, seed=19)
    eig = np.random.default_rng(20).standard_normal((K, B))
    KEref = np.einsum("kbg,kg->kb", np.abs(C) ** 2, ob.KE)
    ref = o_precondition(ob, HC - C * eig[..., None], KEref)
    dev = "cuda"
    Ct = torch.tensor(C, device=dev).to(dtype).contiguous()
    HCt = torch.tensor(HC, device=dev).to(dtype).contiguous()
    out = torch.empty_like(Ct)
    norm2 = torch.empty((K, B), dtype=torch.float64, device=dev)
    rnorm2 = torch.empty_like(norm2)
    eig_t = torch.tensor(eig, device=dev)
    KEref_t = torch.tensor(KEref, device=dev)
    ke_t = torch.tensor(ob.KE, device=dev).contiguous()
    _lib.call("qmx_precond_residual", DTYPE_CODE[dtype], Ct.data_ptr(), HCt.data_ptr(), G, B * G,
              eig_t.data_ptr(), KEref_t.data_ptr(), ke_t.data_ptr(), G, out.data_ptr(), G, B * G,
              K, B, G, norm2.data_ptr(), rnorm2.data_ptr(), rc.stream_ptr())
    tol = 1e-13 if dtype == C128 else 1e-5
    assert rel(out.cpu().numpy(), ref) < tol
    assert rel(norm2.sqrt().cpu().numpy(), np.sqrt((np.abs(ref) ** 2).sum(-1))) < tol
    rn = np.sqrt((np.abs(HC - C * eig[..., None]) ** 2).sum(-1))
    assert rel(rnorm2.sqrt().cpu().numpy(), rn) < (1e-14 if dtype == C128 else 1e-6)
    # KEref formed inside the launch (null pointer): the same residual
    out2 = torch.empty_like(Ct)
    _lib.call("qmx_precond_residual", DTYPE_CODE[dtype], Ct.data_ptr(), HCt.data_ptr(), G, B * G,
              eig_t.data_ptr(), None, ke_t.data_ptr(), G, out2.data_ptr(), G, B * G,
              K, B, G, norm2.data_ptr(), None, rc.stream_ptr())
    assert rel(out2.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("dtype", [C128, C64])
def test_hamiltonian(dtype, small_el, cfg0_el):
    for spec, ob, el in (small_el, cfg0_el):
        C = rand_coeff(ob, 6, seed=8)
        ref = OHam(ob)(C)
        got = el.hamiltonian(to_wf(el.basis, C, dtype)).coeff[0, :, :, 0]
        assert rel(got, ref) < (1e-12 if dtype == C128 else 2e-5)


def test_wavefunction_semantics(cfg0_el):
    spec, ob, el = cfg0_el
    b = el.basis
    A = to_wf(b, rand_coeff(ob, 4, seed=9), C128)
    B = to_wf(b, rand_coeff(ob, 3, seed=10), C128)
    a, bb = A.coeff[0, :, :, 0].cpu().numpy(), B.coeff[0, :, :, 0].cpu().numpy()
    assert rel((A ^ B).wait()[0], np.conj(a) @ bb.transpose(0, 2, 1)) < 1e-13
    M = torch.randn(len(ob.k), 4, 2, dtype=C128, device="cuda")
    assert rel((A @ M).coeff[0, :, :, 0], M.cpu().numpy().transpose(0, 2, 1) @ a) < 1e-13
    # projections: cached, propagated through @ and cat, invalidated by beta_version
    P = A.proj
    assert A._proj_is_valid() and rel(P[0], np.conj(el.ions.beta.coeff[0, :, :, 0].cpu().numpy()) @ a.transpose(0, 2, 1)) < 1e-13
    AM = A @ M
    assert AM._proj_is_valid() and rel(AM._proj, AM.proj.clone()) < 1e-14
    AB = A.cat(B.clone())
    assert AB.n_bands() == 7 and not AB._proj_is_valid()
    A2 = A.clone()
    A2 *= 2.0
    assert A2._proj_is_valid() and rel(A2._proj, 2 * P) < 1e-15
    sl = A[:, :, 1:3]
    assert sl._proj_is_valid() and sl._proj.shape[-1] == 2
    el.ions.beta_version += 1
    assert not A._proj_is_valid()
    el.ions.beta_version -= 1
    O = A.orthonormalize()
    assert rel((O ^ O).wait(), torch.eye(4, dtype=C128, device="cuda").expand(len(ob.k), 4, 4)[None]) < 1e-12


def test_eighg_invariants(gpu):
    n = 40
    A = torch.randn(2, n, n, dtype=C128, device="cuda")
    H = A + A.conj().transpose(-2, -1)
    Bm = torch.randn(2, n, n, dtype=C128, device="cuda")
    O = Bm @ Bm.conj().transpose(-2, -1) + n * torch.eye(n, device="cuda")
    E, V = eighg(H, O)
    assert (H @ V - O @ V * E[:, None, :]).abs().max().item() < 1e-9
    assert (V.conj().transpose(-2, -1) @ O @ V - torch.eye(n, device="cuda")).abs().max().item() < 1e-12


def _solve(spec, dtype, n_iterations, eig_threshold, seed=0, **kw):
    el = electrons_from_spec(spec, dtype=dtype, n_iterations=n_iterations, eig_threshold=eig_threshold, **kw)
    el.initialize_wavefunctions(seed=seed)
    return el, el.solve()


def test_davidson_small_vs_dense(small_el):
    spec, ob, _ = small_el
    exact = dense_eigenvalues(ob, spec.n_bands)
    el, rep = _solve(spec, C128, 60, 1e-11)
    assert rel(rep.eig[0], exact) < 1e-9 / np.abs(exact).max()
    assert rep.residual_norm.max().item() < 1e-4


def test_chefsi_small_vs_dense(small_el):
    spec, ob, _ = small_el
    exact = dense_eigenvalues(ob, spec.n_bands)
    el, rep = _solve(spec, C128, 60, 1e-11, diagonalizer="chefsi", filter_order=8)
    assert np.abs(rep.eig[0].cpu().numpy() - exact).max() < 1e-9


def test_cfg0_plain_matches_oracle_path(cfg0_el):
    """Plain FP64 Davidson on the 2-atom cell: same trajectory as the oracle."""
    spec, ob, _ = cfg0_el
    print("[cfg0 plain] 20 iterations", flush=True)
    res = o_davidson(ob, OHam(ob), o_initial(ob, spec.n_bands), spec.n_bands, spec.n_bands_extra, 20, 1e-8)
    el, rep = _solve(spec, C128, 20, 1e-8)
    assert rep.n_iterations == res.n_iterations
    d = np.abs(rep.eig[0].cpu().numpy() - res.eig[:, : spec.n_bands]).max()
    print(f"[cfg0 plain] max |eig - oracle| = {d:.2e}", flush=True)
    assert d < 1e-8


def test_cfg0_mixed_within_1e6(cfg0_el):
    """Mixed form (complex64 + FP64 Rayleigh-Ritz) within 1e-6 Ha of FP64."""
    spec, ob, _ = cfg0_el
    exact = dense_eigenvalues(ob, spec.n_bands)
    el, rep = _solve(spec, C64, 100, 1e-7)
    d = np.abs(rep.eig[0].cpu().numpy() - exact).max()
    print(f"[cfg0 mixed] iterations {rep.n_iterations}, max |eig - FP64| = {d:.2e}", flush=True)
    assert d < 1e-6


def test_cfg0_chefsi_mixed_residual(cfg0_el):
    """Mixed CheFSI: bands lock and the solve stops on ||HC - eig C|| < 3e-5
    (complex64 filter), then the FP64 Rayleigh-Ritz; within 1e-7 Ha of the
    dense FP64 eigenvalues."""
    spec, ob, _ = cfg0_el
    exact = dense_eigenvalues(ob, spec.n_bands)
    el, rep = _solve(spec, C64, 100, 1e-8, diagonalizer="chefsi", filter_order=8,
                     residual_threshold=3e-5)
    d = np.abs(rep.eig[0].cpu().numpy() - exact).max()
    print(f"[cfg0 chefsi mixed] iterations {rep.n_iterations}, max |eig - FP64| = {d:.2e},"
          f" max residual {rep.residual_norm.max().item():.2e}", flush=True)
    assert rep.n_iterations < 100
    assert d < 1e-7
    assert rep.residual_norm.max().item() < 1e-4


def _hermitian_batch(n, seed):
    """Three random Hermitian matrices plus one with a 4-fold degenerate,
    one zero and one tiny eigenvalue (U diag U^H), all in complex128."""
    g = torch.Generator().manual_seed(seed)
    A = torch.randn(4, n, n, dtype=C128, generator=g)
    H = A + A.conj().transpose(-2, -1)
    lam = torch.linspace(-3.0, 5.0, n, dtype=torch.float64)
    lam[: min(4, n)] = 1.25
    if n > 6:
        lam[5], lam[6] = 0.0, 1e-13
    U, _ = torch.linalg.qr(torch.randn(n, n, dtype=C128, generator=g))
    H[3] = (U * lam) @ U.conj().T
    return H


@pytest.mark.parametrize("method,n", [("jacobi", 1), ("jacobi", 36), ("jacobi", 96),
                                      ("ql", 1), ("ql", 2), ("ql", 37), ("ql", 96), ("ql", 128),
                                      ("ql", 142), ("ql", 200)])
def test_small_eigensolvers(gpu, method, n):
    """Batched on-chip eigensolvers: eigenvalues equal LAPACK's (numpy, FP64),
    eigenvectors are orthonormal and satisfy H V = V E."""
    from qimpy_amd.linalg import eigh_small
    H = _hermitian_batch(n, seed=n)
    E, V = eigh_small(H.cuda(), method)
    if method == "ql":
        assert int(eigh_small.last_info.abs().sum().item()) == 0
    E, V = E.cpu(), V.cpu()
    ref = np.linalg.eigvalsh(H.numpy())
    scale = np.abs(ref).max()
    assert np.abs(E.numpy() - ref).max() < 1e-12 * scale
    I = torch.eye(n, dtype=C128)
    assert (V.conj().transpose(-2, -1) @ V - I).abs().max().item() < 1e-12
    assert (H @ V - V * E[:, None, :]).abs().max().item() < 1e-11 * scale


@pytest.mark.parametrize("shape", [(8, 36), (8, 96), (1, 300)])
def test_eigh_dispatch(gpu, shape):
    """linalg.eigh routes (batch, n) to Jacobi / QL / cooperative tridiagonal
    reduction; every route agrees with LAPACK to 1e-12 relative."""
    from qimpy_amd.linalg import eigh
    B, n = shape
    g = torch.Generator().manual_seed(B * 1000 + n)
    A = torch.randn(B, n, n, dtype=C128, generator=g)
    H = A + A.conj().transpose(-2, -1)
    E, V = eigh(H.cuda())
    E, V = E.cpu(), V.cpu()
    ref = np.linalg.eigvalsh(H.numpy())
    scale = np.abs(ref).max()
    assert np.abs(E.numpy() - ref).max() < 1e-12 * scale
    assert (H @ V - V * E[:, None, :]).abs().max().item() < 1e-11 * scale


def test_ql_large_batch_global_accumulator(gpu):
    """n > 128 keeps the QL accumulator in the workspace, one region per
    matrix: a batch of 48 at n = 144 (the largest subspace of 72-band
    cells) must give every matrix's own eigenpairs."""
    from qimpy_amd.linalg import eigh_small
    n, B = 144, 48
    g = torch.Generator().manual_seed(144)
    A = torch.randn(B, n, n, dtype=C128, generator=g)
    H = A + A.conj().transpose(-2, -1)
    E, V = eigh_small(H.cuda(), "ql")
    assert int(eigh_small.last_info.abs().sum().item()) == 0
    E, V = E.cpu(), V.cpu()
    ref = np.linalg.eigvalsh(H.numpy())
    scale = np.abs(ref).max()
    assert np.abs(E.numpy() - ref).max() < 1e-12 * scale
    assert (H @ V - V * E[:, None, :]).abs().max().item() < 1e-11 * scale


def test_eigh_grouped_mixed_sizes(gpu):
    """One grouped launch over batches of different n (both accumulator
    placements): every matrix matches LAPACK; an empty batch is allowed."""
    from qimpy_amd.linalg import eigh_grouped
    g = torch.Generator().manual_seed(7)
    mats = []
    for B, n in ((3, 5), (2, 37), (4, 128), (2, 144), (0, 20), (1, 1)):
        A = torch.randn(B, n, n, dtype=C128, generator=g)
        mats.append(A + A.conj().transpose(-2, -1))
    results, info = eigh_grouped([H.cuda() for H in mats])
    assert int(info.abs().sum().item()) == 0
    for H, (E, V) in zip(mats, results):
        if H.shape[0] == 0:
            assert E.shape == (0, H.shape[-1])
            continue
        E, V = E.cpu(), V.cpu()
        ref = np.linalg.eigvalsh(H.numpy())
        scale = np.abs(ref).max()
        assert np.abs(E.numpy() - ref).max() < 1e-12 * scale
        assert (H @ V - V * E[:, None, :]).abs().max().item() < 1e-11 * scale


def _unitary(n, rng):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    return Q


def _bisect_cases():
    """(name, H (1, n, n) complex128, n_want) for the bisection eigensolver:
    random spectra, exactly degenerate clusters (multiplicity 2, 3, 8, 16),
    decoupled rows shifted to 1e6 (the dropped directions of the batched
    Davidson step), the zero matrix, the identity and a diagonal matrix."""
    rng = np.random.default_rng(11)
    out = []
    for n, w in ((1, 1), (2, 2), (3, 1), (36, 36), (64, 32), (144, 72), (200, 200), (256, 40)):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        out.append((f"random{n}", A + A.conj().T, w))
    for n, mult in ((12, 3), (72, 8), (144, 2), (144, 3), (144, 16)):
        ev = np.repeat(rng.standard_normal((n + mult - 1) // mult), mult)[:n]
        Q = _unitary(n, rng)
        out.append((f"degenerate{n}x{mult}", (Q * ev) @ Q.conj().T, n // 2))
    A = rng.standard_normal((144, 144)) + 1j * rng.standard_normal((144, 144))
    H = A + A.conj().T
    H[134:, :] = 0
    H[:, 134:] = 0
    H[np.arange(134, 144), np.arange(134, 144)] = 1e6
    out.append(("shifted", H, 72))
    out.append(("zero", np.zeros((10, 10), complex), 10))
    out.append(("identity", np.eye(20, dtype=complex), 20))
    out.append(("diagonal", np.diag(np.arange(30.0)).astype(complex), 15))
    return [(name, torch.tensor(H)[None], w) for name, H, w in out]


@pytest.mark.parametrize("case", _bisect_cases(), ids=lambda c: c[0])
def test_eigh_bisect_cases(gpu, case):
    """Bisection + inverse-iteration eigensolver (qmx_heevx_grouped) against
    LAPACK: all eigenvalues to 1e-13 of the spectral radius; the n_want
    lowest eigenvectors with residual |HV - VE| < 1e-12 |H| and
    orthonormality to 1e-11 (inverse iteration leaves about eps |H| / gap
    between neighbours outside a cluster, removed inside clusters)."""
    from qimpy_amd.linalg import eigh_grouped
    name, H, w = case
    (res,), info = eigh_grouped([H.cuda()], [w])
    E, V = res
    assert int(info.abs().sum().item()) == 0
    E, V = E.cpu()[0], V.cpu()[0, :, :w]
    ref = np.linalg.eigvalsh(H[0].numpy())
    scale = max(np.abs(ref).max(), 1e-300)
    assert np.abs(E.numpy() - ref).max() <= 1e-13 * scale
    assert (H[0] @ V - V * E[None, :w]).abs().max().item() <= 1e-12 * scale
    assert (V.conj().T @ V - torch.eye(w, dtype=C128)).abs().max().item() < 1e-11


def test_eigh_bisect_decoupled_shift(gpu):
    """Rows decoupled at 1e6 (the dropped directions of a Davidson step) do
    not degrade the wanted spectrum below them: eigenvalues, residuals and
    orthonormality of the 72 lowest eigenvectors are checked against the
    norm of the coupled block (about 30), which holds close pairs (gaps of
    1e-7 and 1e-4) and one exactly degenerate pair."""
    from qimpy_amd.linalg import eigh_grouped
    rng = np.random.default_rng(5)
    nb, n = 134, 144
    ev = np.sort(rng.standard_normal(nb)) * 10.0
    ev[1], ev[11], ev[21] = ev[0], ev[10] + 1e-7, ev[20] + 1e-4
    Q = _unitary(nb, rng)
    H = np.zeros((n, n), complex)
    H[:nb, :nb] = (Q * ev) @ Q.conj().T
    H[np.arange(nb, n), np.arange(nb, n)] = 1e6
    w = 72
    (res,), info = eigh_grouped([torch.tensor(H)[None].cuda()], [w])
    assert int(info.abs().sum().item()) == 0
    E, V = res[0].cpu()[0], res[1].cpu()[0, :, :w]
    ref = np.linalg.eigvalsh(H)
    scale = np.abs(ev).max()
    assert np.abs(E.numpy()[:w] - ref[:w]).max() <= 1e-9 * scale
    Ht = torch.tensor(H)
    assert (Ht @ V - V * E[None, :w]).abs().max().item() <= 1e-9 * scale
    assert (V.conj().T @ V - torch.eye(w, dtype=C128)).abs().max().item() < 1e-11
    assert V[nb:].abs().max().item() < 1e-12  # no weight on the decoupled rows


def test_eigh_bisect_nonfinite(gpu):
    """A NaN in the input gives info = 1 and NaN eigenvalues, no hang."""
    from qimpy_amd.linalg import eigh_grouped
    H = torch.eye(8, dtype=C128)[None].clone()
    H[0, 3, 2] = float("nan")
    (res,), info = eigh_grouped([H.cuda()])
    assert int(info[0].item()) == 1
    assert torch.isnan(res[0]).all()


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("shape", [(1, 1, 1, 1), (3, 5, 7, 13), (2, 70, 33, 1000), (4, 144, 72, 4017),
                                   (1, 130, 65, 40000)])
def test_dot128_matches_fp64_reference(gpu, dtype, shape):
    """Split-K MFMA overlap product conj(A) B^T against an FP64 CPU product:
    complex128 operands to FP64 rounding, complex64 operands to FP32
    accumulation of chunks (relative 1e-5 of |A||B|); also on a strided
    batch (rows inside a larger buffer) and with accumulate."""
    from qimpy_amd.gemm import dot128
    batch, M, N, G = shape
    g = torch.Generator().manual_seed(M * 7 + N)
    A = torch.randn(batch, M + N, G, dtype=C128, generator=g)  # A rows, then B rows
    ref = torch.matmul(A[:, :M].conj(), A[:, M:].transpose(-2, -1))
    Ad = A.to(dtype).cuda()
    got = dot128(Ad[:, :M], Ad[:, M:])  # strided views of one buffer
    scale = float(A[:, :M].abs().max() * A[:, M:].abs().max() * G)
    tol = 1e-12 if dtype == C128 else 1e-6
    assert got.dtype == C128
    assert (got.cpu() - ref).abs().max().item() <= tol * scale
    again = dot128(Ad[:, :M], Ad[:, M:], out=got.clone(), accumulate=True)
    assert (again.cpu() - 2 * ref).abs().max().item() <= 2 * tol * scale
    lazy = dot128(Ad[:, :M].conj(), Ad[:, M:])  # torch's lazy conjugate view
    ref_c = torch.matmul(A[:, :M], A[:, M:].transpose(-2, -1))
    assert (lazy.cpu() - ref_c).abs().max().item() <= tol * scale


def _radial_samples(l_rf, n_G, dG, seed):
    """Smooth radial functions (Gaussian times cosine) sampled on the nodes."""
    rng = np.random.default_rng(seed)
    G = np.arange(n_G) * dG
    return np.stack([np.exp(-0.5 * a * G**2) * np.cos(b * G)
                     for a, b in rng.uniform([0.5, 0.2], [2.0, 1.5], (len(l_rf), 2))])


@pytest.mark.parametrize("l_rf", [(0, 1, 2, 3, 2), (4, 6), (5,)])
def test_spline_projectors_vs_oracle(gpu, l_rf):
    """qmx_projectors (spline radials, real solid harmonics up to l = 6,
    translation phases, padding) against the oracle's projectors built from
    scipy's harmonics and B-splines on the same blip coefficients: 1e-12
    of the largest entry."""
    from oracle.projectors import projectors as o_spline_projectors
    from oracle.projectors import quintic_coeff as o_coeff
    from qimpy_amd.ions import RadialSet, spline_projectors

    spec = _cell_with_box(24)
    print(f"[spline projectors] l {l_rf}", flush=True)
    ob = OBasis(spec)
    el = electrons_from_spec(spec)
    dG = 0.02
    n_G = int(np.ceil(np.sqrt(2 * spec.ke_cutoff) / dG)) + 10
    f = _radial_samples(l_rf, n_G, dG, seed=len(l_rf))
    radial = RadialSet(l_rf, f, dG)
    got = spline_projectors(el.basis, torch.tensor(spec.positions), radial).cpu().numpy()
    want = o_spline_projectors(ob, spec.positions, o_coeff(f), dG, l_rf, spec.volume)
    assert got.shape == want.shape
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"[spline projectors] max err {err:.2e}", flush=True)
    assert err < 1e-12
    pad = np.broadcast_to(ob.pad_mask[:, None, :], got.shape)
    assert np.all(got[pad] == 0)


def test_spline_ions_reproduce_gaussian_projectors(gpu):
    """Ions with a RadialSet of sampled Gaussians (spacing 0.02) give the
    analytic Gaussian projectors of the synthetic cell to 1e-8, and the
    library refuses tables short of the sphere and more than 32 projectors."""
    from qimpy_amd.ions import Ions, RadialSet, spline_projectors

    spec = _cell_with_box(30)
    el = electrons_from_spec(spec)
    b = el.basis
    dG = 0.02
    G = np.arange(int(np.ceil(np.sqrt(2 * spec.ke_cutoff) / dG)) + 10) * dG
    l_rf = [l for rc in spec.proj_rc for l in spec.proj_l]
    f = np.stack([np.exp(-0.5 * rc**2 * G**2) for rc in spec.proj_rc for _ in spec.proj_l])
    ions = Ions(spec.positions, (), (), spec.D, radial=RadialSet(l_rf, f, dG))
    ions.update(b)
    want = el.ions.beta.coeff
    assert ions.n_projectors == el.ions.n_projectors
    err = float((ions.beta.coeff - want).abs().max() / want.abs().max())
    print(f"[spline ions] max err vs Gaussian {err:.2e}", flush=True)
    assert err < 1e-8
    with pytest.raises(ValueError, match="sampled to"):
        spline_projectors(b, torch.tensor(spec.positions), RadialSet(l_rf, f[:, :20], dG))
    with pytest.raises(RuntimeError, match="more than 32"):
        spline_projectors(b, torch.tensor(spec.positions),
                          RadialSet((3,) * 5, np.repeat(f[:1], 5, axis=0), dG))


@pytest.mark.parametrize("dtype", [C128, C64])
def test_gradient_and_tau_branch(cfg0_el, dtype):
    """qmx_apply_gradient (i (G + k)_dir C) and the meta-GGA term of H,
    -1/2 sum_dir grad(V_tau grad C), against the oracle: complex128 to
    1e-12, complex64 to 2e-5 of the largest entry."""
    from oracle.hamiltonian import apply_gradient as o_grad
    from oracle.hamiltonian import apply_tau as o_tau
    spec, ob, el = cfg0_el
    b = el.basis
    C = rand_coeff(ob, 5, seed=12)
    wf = to_wf(b, C, dtype)
    tol = 1e-12 if dtype == C128 else 2e-5
    for d in range(3):
        assert rel(b.apply_gradient(wf, d).coeff[0, :, :, 0], o_grad(ob, C, d)) < tol
    V_tau = np.random.default_rng(3).uniform(0.2, 1.0, ob.shape)
    out = torch.zeros_like(wf.coeff)
    b.apply_tau(torch.tensor(V_tau, device="cuda"), wf.coeff, out)
    err = rel(out[0, :, :, 0], o_tau(ob, V_tau, C))
    print(f"[tau branch] {dtype} {err:.2e}", flush=True)
    assert err < tol


def test_davidson_with_tau_vs_dense(small_el):
    """A band solve whose H carries the meta-GGA term reaches the dense
    eigenvalues of T + V + 1/2 (G + k) . (G' + k) V~tau(G - G') + nonlocal."""
    spec, ob, _ = small_el
    V_tau = np.random.default_rng(4).uniform(0.1, 0.5, ob.shape)
    exact = dense_eigenvalues(ob, spec.n_bands, V_tau=V_tau)
    el, rep = _solve(spec, C128, 80, 1e-11, V_tau=V_tau)
    d = np.abs(rep.eig[0].cpu().numpy() - exact).max()
    print(f"[davidson tau] iterations {rep.n_iterations}, max |eig - dense| {d:.2e}", flush=True)
    assert d < 1e-9


def test_radial_transform_vs_oracle(gpu):
    """qmx_radial_transform (own j_l(x)/x^l: Taylor series and upward
    recurrence) against the oracle's scipy spherical_jn for l = 0..6 on a
    non-uniform radial grid: 1e-11 of the largest value per function."""
    from oracle.projectors import radial_transform as o_transform
    from qimpy_amd.ions import radial_transform

    t = np.linspace(0.0, 1.0, 1201)
    r = 10.0 * t**2 + 2.0 * t  # denser near the origin
    dr = (20.0 * t + 2.0) * (t[1] - t[0])
    l_rf = [0, 1, 2, 3, 4, 5, 6]
    f = np.stack([np.exp(-0.4 * r**2) * (1.0 + 0.3 * np.cos((l + 1) * r)) for l in l_rf])
    got = radial_transform(r, dr, f, l_rf, 9.0, 0.02)
    want = o_transform(r, dr, f, l_rf, 9.0, 0.02)
    assert got.shape == want.shape
    for i in range(len(l_rf)):
        err = np.abs(got[i] - want[i]).max() / np.abs(want[i]).max()
        print(f"[radial transform] l {l_rf[i]}: {err:.2e}", flush=True)
        assert err < 1e-11


def test_projectors_from_real_space_gaussians(gpu):
    """The whole chain on the device: real-space Gaussians (2 pi)^(-3/2)
    rc^-(2l+3) exp(-r^2 / 2 rc^2) -> qmx_radial_transform -> blip spline ->
    qmx_projectors equals the analytic projectors exp(-rc^2 G^2 / 2) of the
    synthetic cell to 1e-8."""
    from qimpy_amd.ions import Ions, RadialSet

    spec = _cell_with_box(30)
    el = electrons_from_spec(spec)
    r = np.linspace(0.0, 12 * max(spec.proj_rc), 3001)
    dr = np.full_like(r, r[1] - r[0])
    l_rf, f = [], []
    for rc_ in spec.proj_rc:
        for l in spec.proj_l:
            l_rf.append(l)
            f.append((2 * np.pi) ** -1.5 * rc_ ** -(2 * l + 3) * np.exp(-0.5 * r**2 / rc_**2))
    G_max = np.sqrt(2 * spec.ke_cutoff)
    radial = RadialSet.from_real_space(l_rf, r, dr, np.stack(f), G_max)
    ions = Ions(spec.positions, (), (), spec.D, radial=radial)
    ions.update(el.basis)
    want = el.ions.beta.coeff
    err = float((ions.beta.coeff - want).abs().max() / want.abs().max())
    print(f"[real-space gaussians] max err {err:.2e}", flush=True)
    assert err < 1e-8


@pytest.mark.parametrize("l_rf", [(0,), (0, 1, 1)])
def test_lcao_start_vs_dense(small_el, l_rf):
    """LCAO start (atomic orbitals from qmx_projectors, the rest random, FP64
    subspace eighg): orthonormal bands whose Ritz values bound the dense
    eigenvalues from above (variational), then a Davidson solve from it
    reaches the dense eigenvalues. l = (0,) gives fewer orbitals than bands,
    (0, 1, 1) more."""
    from qimpy_amd.ions import RadialSet
    spec, ob, _ = small_el
    exact = dense_eigenvalues(ob, spec.n_bands)
    el = electrons_from_spec(spec, dtype=C128, n_iterations=80, eig_threshold=1e-11)
    G = np.arange(int(np.ceil(np.sqrt(2 * spec.ke_cutoff) / 0.02)) + 10) * 0.02
    f = np.stack([np.exp(-0.5 * (0.8 + 0.4 * i) ** 2 * G**2) for i in range(len(l_rf))])
    el.initialize_lcao(RadialSet(l_rf, f))
    n_total = el.C.n_bands()
    O = (el.C ^ el.C).wait()
    assert (O - torch.eye(n_total, device="cuda")).abs().max().item() < 1e-10
    eig0 = el.eig[0].cpu().numpy()
    assert np.all(np.diff(eig0, axis=-1) >= -1e-12)
    assert np.all(eig0[:, : spec.n_bands] >= exact - 1e-10)
    rep = el.solve()
    d = np.abs(rep.eig[0].cpu().numpy() - exact).max()
    print(f"[lcao {l_rf}] start max {np.abs(eig0[:, :spec.n_bands] - exact).max():.2e},"
          f" iterations {rep.n_iterations}, max |eig - dense| {d:.2e}", flush=True)
    assert d < 1e-9


def test_cfg4_chefsi_full_size(gpu):
    """The 256-atom CheFSI solve at full size (144^3 box, 1024 bands, filter
    order 20, ten outer iterations at most), by size-independent properties:
    both forms finish within the iteration limit, their eigenvalues are
    sorted, and the mixed form (complex64 filter, residual 3e-5, FP64
    Rayleigh-Ritz) is within 1e-7 Ha of the plain FP64 solve everywhere."""
    spec = systems.make_cell(4)
    out = {}
    for dt, kw in ((C128, {}), (C64, {"residual_threshold": 3e-5})):
        print(f"[cfg4 chefsi] {dt}", flush=True)
        el = electrons_from_spec(spec, dtype=dt, **kw)
        el.initialize_wavefunctions(seed=0)
        rep = el.solve()
        eig = rep.eig.cpu().numpy()
        print(f"[cfg4 chefsi] {dt}: {rep.n_iterations} outer iterations", flush=True)
        assert rep.n_iterations <= spec.n_iterations
        assert np.all(np.diff(eig, axis=-1) >= -1e-10)
        out[dt] = eig
        del el, rep
        torch.cuda.empty_cache()
    d = np.abs(out[C64] - out[C128]).max()
    print(f"[cfg4 chefsi] max |eig mixed - plain| {d:.2e}", flush=True)
    assert d < 1e-7
