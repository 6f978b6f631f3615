"""Pin the FP64 oracle against closed forms, brute force and library routines.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails it.
"""
import itertools

import numpy as np
import pytest
import scipy.linalg

from oracle.basis import Basis
from oracle.davidson import chefsi, davidson, e_band, initial_state
from oracle.dense import dense_eigenvalues, dense_hamiltonian
from oracle.hamiltonian import (Hamiltonian, apply_nonlocal, apply_potential, dot,
                                matmul, projectors)
from oracle.init import randomize, xorshift_draw
from oracle.linalg import eighg
from qimpy_amd import systems


@pytest.fixture(scope="module")
def small():
    spec = systems.small_cell(n_bands=5)
    return spec, Basis(spec)


def test_basis_brute_force(small):
    """Sphere counts and fft_index order from an explicit enumeration."""
    spec, b = small
    n0, n1, n2 = spec.shape
    for ik, k in enumerate(spec.k):
        pts = []
        for i, j, l in itertools.product(range(n0), range(n1), range(n2)):
            f = np.array([i if i <= n0 // 2 else i - n0, j if j <= n1 // 2 else j - n1,
                          l if l <= n2 // 2 else l - n2])
            G = (f + k) @ spec.Gbasis.T
            if 0.5 * G @ G < spec.ke_cutoff:
                pts.append((i * n1 + j) * n2 + l)
        assert b.n[ik] == len(pts)
        assert np.array_equal(b.fft_index[ik, : b.n[ik]], np.array(sorted(pts)))
        # padding: the first out-of-sphere indices, ascending
        out = sorted(set(range(n0 * n1 * n2)) - set(pts))[: b.n_tot - b.n[ik]]
        assert np.array_equal(b.fft_index[ik, b.n[ik]:], np.array(out, dtype=int))


def test_workload_sizes():
    """Basis sizes of the measured workloads (fixes the per-unit figures)."""
    b0 = Basis(systems.make_cell(0))
    assert (b0.n.min(), b0.n.max()) == (1150, 1152)
    b1 = Basis(systems.make_cell(1))
    assert b1.n_tot == 16135


def test_fft_convention_direct_dft():
    """norm='forward' fft = (1/N) sum exp(-iGr); ifft unscaled, on a 3x4x5 box."""
    import scipy.fft

    rng = np.random.default_rng(1)
    shape = (3, 4, 5)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    r = np.stack(np.meshgrid(*[np.arange(s) for s in shape], indexing="ij"), -1).reshape(-1, 3)
    phase = np.exp(-2j * np.pi * (r[:, None, :] / shape * r[None, :, :]).sum(-1))
    direct = (phase @ x.reshape(-1)) / x.size  # G index on axis 0
    got = scipy.fft.fftn(x, norm="forward").reshape(-1)
    assert np.abs(got - direct).max() < 1e-13
    inv = scipy.fft.ifftn(x, norm="forward").reshape(-1)
    assert np.abs(inv - phase.conj() @ x.reshape(-1)).max() < 1e-12
    # Parseval with the reference's normalisation (grid/test_field.py:14-22):
    xt = scipy.fft.fftn(x, norm="forward")
    assert abs(np.vdot(x, x) / x.size - np.vdot(xt, xt)) < 1e-12 * np.vdot(x, x).real


def test_apply_potential_convolution(small):
    """(V psi)(G) = sum_G' V~(G - G') psi(G') by explicit double sum."""
    spec, b = small
    rng = np.random.default_rng(2)
    C = rng.standard_normal((len(spec.k), 2, b.n_tot)) + 1j * rng.standard_normal((len(spec.k), 2, b.n_tot))
    b.constrain(C)
    got = apply_potential(b, spec.V, C)
    Vt = np.fft.fftn(spec.V) / b.N
    for ik in range(len(spec.k)):
        n = b.n[ik]
        iG = b.iG[ik, :n]
        diff = (iG[:, None] - iG[None, :]) % np.array(spec.shape)
        M = Vt[diff[..., 0], diff[..., 1], diff[..., 2]]
        assert np.abs(got[ik, :, :n] - C[ik, :, :n] @ M.T).max() < 1e-13
        if n < b.n_tot:
            assert np.abs(got[ik, :, n:]).max() == 0.0


def test_apply_potential_constant_and_hermitian(small):
    spec, b = small
    rng = np.random.default_rng(3)
    C = rng.standard_normal((len(spec.k), 3, b.n_tot)) + 1j * rng.standard_normal((len(spec.k), 3, b.n_tot))
    b.constrain(C)
    out = apply_potential(b, np.full(spec.shape, 0.37), C)
    assert np.abs(out - 0.37 * C).max() < 1e-14
    VC = apply_potential(b, spec.V, C)
    lhs = dot(C[:, :1], VC[:, 1:])
    rhs = dot(VC[:, :1], C[:, 1:])
    assert np.abs(lhs - rhs).max() < 1e-13


def test_projectors_closed_form(small):
    """beta for one atom at the origin: s-channel is exp(-rc^2 G^2/2)/sqrt(4 pi V),
    p-channel is i sqrt(3/4pi) (y, z, x) exp(...)/sqrt(V)."""
    spec, b = small
    spec0 = systems.small_cell(n_atoms=1)
    spec0.positions[:] = 0.0
    b0 = Basis(spec0)
    beta = projectors(b0)
    Gk = (b0.iG + b0.k[:, None]) @ b0.Gbasis.T
    G2 = (Gk**2).sum(-1)
    f = np.exp(-0.5 * 0.8**2 * G2) / np.sqrt(spec0.volume)
    mask = ~b0.pad_mask
    assert np.abs((beta[:, 0] - f / np.sqrt(4 * np.pi))[mask]).max() < 1e-14
    py = 1j * np.sqrt(3 / (4 * np.pi)) * Gk[..., 1] * f
    assert np.abs((beta[:, 1] - py)[mask]).max() < 1e-14
    assert np.abs(beta[:, 3][mask] - (1j * np.sqrt(3 / (4 * np.pi)) * Gk[..., 0] * f)[mask]).max() < 1e-14
    # translation: moving the atom multiplies by exp(-2 pi i (iG + k).tau)
    spec0.positions[:] = [[0.1, 0.2, 0.3]]
    beta_t = projectors(b0)
    ph = np.exp(-2j * np.pi * ((b0.iG + b0.k[:, None]) @ np.array([0.1, 0.2, 0.3])))
    assert np.abs((beta_t[:, 0] - beta[:, 0] * ph)[mask]).max() < 1e-14


def test_nonlocal_dense(small):
    """beta D beta^H C equals the explicit dense sum and is Hermitian."""
    spec, b = small
    beta = projectors(b)
    rng = np.random.default_rng(4)
    C = rng.standard_normal((len(spec.k), 3, b.n_tot)) + 1j * rng.standard_normal((len(spec.k), 3, b.n_tot))
    b.constrain(C)
    got = apply_nonlocal(beta, spec.D, C)
    for ik in range(len(spec.k)):
        Hnl = (beta[ik].T * spec.D) @ beta[ik].conj()
        assert np.abs(got[ik] - C[ik] @ Hnl.T).max() < 1e-13
        assert np.abs(Hnl - Hnl.conj().T).max() < 1e-14


def test_dense_matches_operator(small):
    """The dense H (independent of the FFT) applied to C equals the H*psi chain."""
    spec, b = small
    H = Hamiltonian(b)
    rng = np.random.default_rng(5)
    C = rng.standard_normal((len(spec.k), 2, b.n_tot)) + 1j * rng.standard_normal((len(spec.k), 2, b.n_tot))
    b.constrain(C)
    HC = H(C)
    for ik in range(len(spec.k)):
        n = b.n[ik]
        Hd = dense_hamiltonian(b, ik, beta=H.beta)
        assert np.abs(HC[ik, :, :n] - C[ik, :, :n] @ Hd.T).max() < 1e-12
        assert np.abs(Hd - Hd.conj().T).max() < 1e-13


def test_eighg_invariants():
    rng = np.random.default_rng(6)
    n = 12
    A = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    H = A + A.conj().transpose(0, 2, 1)
    B = rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n))
    O = B @ B.conj().transpose(0, 2, 1) + n * np.eye(n)
    E, V = eighg(H, O)
    for i in range(2):
        assert np.abs(H[i] @ V[i] - O[i] @ V[i] * E[i]).max() < 1e-10
        assert np.abs(V[i].conj().T @ O[i] @ V[i] - np.eye(n)).max() < 1e-12
        assert np.all(np.diff(E[i]) >= 0)
        assert np.abs(E[i] - scipy.linalg.eigh(H[i], O[i], eigvals_only=True)).max() < 1e-10


def test_xorshift_bigint():
    """Signed int64 xorshift (arithmetic >>) against Python big integers."""
    seeds = np.array([1, 2, 12345, -7, 2**62 + 3, -(2**63) + 5], dtype=np.int64)
    x = seeds.copy()

    def ref_step(v):
        m = (1 << 64) - 1

        def s64(u):
            return u - (1 << 64) if u >= 1 << 63 else u

        v = s64((v ^ ((v << 13) & m)) & m)
        v = s64((v ^ (v >> 7)) & m)  # Python >> on negative ints is arithmetic
        v = s64((v ^ ((v << 17) & m)) & m)
        return v

    vals = [int(s) for s in seeds]
    for _ in range(5):
        with np.errstate(over="ignore"):
            u = xorshift_draw(x)
        vals = [ref_step(v) for v in vals]
        assert [int(a) for a in x] == vals
        assert np.all(u == np.array([(0.5**64) * float(v) + 0.5 for v in vals]))


def test_randomize_window_consistency(small):
    """Randomizing a band window equals the same bands of a full draw
    (the property of the reference's test_random_equiv)."""
    spec, b = small
    full = np.zeros((len(spec.k), 8, b.n_tot), complex)
    randomize(b, full, seed=3)
    part = np.zeros_like(full)
    randomize(b, part, seed=3, b_start=2, b_stop=6)
    assert np.array_equal(part[:, 2:6], full[:, 2:6])
    assert np.abs(part[:, :2]).max() == 0 and np.abs(part[:, 6:]).max() == 0
    assert np.abs(full[np.broadcast_to(b.pad_mask[:, None], full.shape)]).max() == 0


def test_orthonormalize(small):
    """C^H C = I after the Cholesky orthonormalisation (test_orthonormalize)."""
    spec, b = small
    C = initial_state(b, 6)
    assert np.abs(dot(C, C) - np.eye(6)[None]).max() < 1e-12


def test_davidson_and_chefsi_vs_dense(small):
    spec, b = small
    H = Hamiltonian(b)
    exact = dense_eigenvalues(b, spec.n_bands)
    C = initial_state(b, spec.n_bands)
    r = davidson(b, H, C, spec.n_bands, spec.n_bands_extra, 60, 1e-11)
    assert r.converged
    assert np.abs(r.eig[:, : spec.n_bands] - exact).max() < 1e-9
    r2 = chefsi(b, H, C, spec.n_bands, spec.n_bands_extra, spec.V.max(), 8, 60, 1e-11)
    assert r2.converged
    assert np.abs(r2.eig[:, : spec.n_bands] - exact).max() < 1e-9
    assert abs(e_band(b, r.eig, spec.n_bands) - (2 * spec.wk[:, None] * exact).sum()) < 1e-8


def test_chebyshev_recurrence_diagonal():
    """The scaled recurrence of the filter (sigma = e / (c - a), _chefsi.py:122-140)
    equals T_m((x - c)/e) / T_m((c - a)/e) on a diagonal H."""
    x = np.linspace(-2, 5, 41)
    a, lo, up, m = -2.0, 0.5, 5.0, 7
    e, c = 0.5 * (up - lo), 0.5 * (up + lo)
    sigma = e / (c - a)
    tau = 2 / sigma
    C = np.ones_like(x)
    Y = (x - c) * C * sigma / e
    for _ in range(m - 1):
        sn = 1 / (tau - sigma)
        Yt = (x * Y - c * Y) * (2 * sn / e) - (sigma * sn) * C
        C, Y, sigma = Y, Yt, sn
    cheb = lambda t: np.polynomial.chebyshev.chebval(t, [0] * m + [1])
    expect = cheb((x - c) / e) / cheb((c - a) / e)
    assert np.abs(Y - expect).max() < 1e-10 * np.abs(expect).max()
    # inside [lo, up] the filter is bounded by 1/|T_m((c - a)/e)|
    inside = (x >= lo) & (x <= up)
    assert np.abs(Y[inside]).max() <= 1.0 / abs(cheb((c - a) / e)) * (1 + 1e-12)


def test_density_closed_forms(small):
    """Oracle density: one plane wave gives f w |c|^2 / V everywhere; two give
    |a|^2 + |b|^2 + 2 Re(a b* exp(2 pi i (m1 - m2) . r / S)) point by point
    (m the integer box frequencies); normalised random bands integrate to
    sum f w over the cell; all against a direct O(N^2) inverse DFT."""
    from oracle.density import collect_density

    spec, b = small
    K, G = b.fft_index.shape
    rng = np.random.default_rng(3)
    shape = np.array(b.shape)
    r = np.stack(np.meshgrid(*[np.arange(s) for s in b.shape], indexing="ij"), -1)
    # one plane wave in band 0 of k 0, filling 1 (other bands and k empty)
    C = np.zeros((K, 2, G), complex)
    C[0, 0, 3] = 0.6 - 0.8j
    f = np.zeros((K, 2))
    f[0, 0] = 1.0
    rho = collect_density(b, C, f)
    assert np.abs(rho - b.w_sk[0] / b.volume).max() < 1e-13 * b.w_sk[0] / b.volume
    # two plane waves (a at sphere point 1, c at 5) in band 1 of k 1, filling 0.5
    a, c = 0.3 + 0.4j, -0.7 + 0.2j
    C[:] = 0.0
    C[1, 1, 1], C[1, 1, 5] = a, c
    f[:] = 0.0
    f[1, 1] = 0.5
    m1, m2 = b.iG[1, 1], b.iG[1, 5]
    phase = np.exp(2j * np.pi * ((m1 - m2) * r / shape).sum(-1))
    want = 0.5 * b.w_sk[1] / b.volume * (abs(a) ** 2 + abs(c) ** 2 + 2 * (a * np.conj(c) * phase).real)
    assert np.abs(collect_density(b, C, f) - want).max() < 1e-13
    # normalised random bands: cell integral (rho dV, dV = V / N) = sum f w
    C = rng.standard_normal((K, 3, G)) + 1j * rng.standard_normal((K, 3, G))
    b.constrain(C)
    C /= np.linalg.norm(C, axis=-1, keepdims=True)
    f = rng.uniform(0.0, 1.0, (K, 3))
    f[0, 2] = 0.0
    rho = collect_density(b, C, f)
    assert abs(rho.sum() * b.volume / b.N - (f * b.w_sk[:, None]).sum()) < 1e-12
    # direct inverse DFT of k 0, band 1
    rr = r.reshape(-1, 3)
    psi = np.exp(2j * np.pi * ((rr[:, None, :] * b.iG[0][None]) / shape).sum(-1)) @ C[0, 1]
    f1 = np.zeros((K, 3))
    f1[0, 1] = 1.0
    want = b.w_sk[0] / b.volume * np.abs(psi) ** 2
    assert np.abs(collect_density(b, C, f1).reshape(-1) - want).max() < 1e-12 * want.max()


def test_density_cross_identities(small):
    """Oracle spinor cross term X = sum f w / V psi_1 conj(psi_0): the
    polarization identities rho(C0 + C1) - rho(C0) - rho(C1) = 2 Re X and
    rho(C0 + i C1) - rho(C0) - rho(C1) = -2 Im X against the pinned density;
    plane waves a (component 0) and c (component 1) give
    c conj(a) exp(2 pi i (m1 - m0) . r / S)."""
    from oracle.density import collect_density, collect_density_cross

    spec, b = small
    K, G = b.fft_index.shape
    rng = np.random.default_rng(4)
    C0 = rng.standard_normal((K, 3, G)) + 1j * rng.standard_normal((K, 3, G))
    C1 = rng.standard_normal((K, 3, G)) + 1j * rng.standard_normal((K, 3, G))
    b.constrain(C0)
    b.constrain(C1)
    f = rng.uniform(0.0, 1.0, (K, 3))
    f[1, 0] = 0.0
    X = collect_density_cross(b, C0, C1, f)
    r0, r1 = collect_density(b, C0, f), collect_density(b, C1, f)
    scale = np.abs(r0).max()
    assert np.abs(collect_density(b, C0 + C1, f) - r0 - r1 - 2 * X.real).max() < 1e-12 * scale
    assert np.abs(collect_density(b, C0 + 1j * C1, f) - r0 - r1 + 2 * X.imag).max() < 1e-12 * scale
    # swapping the components conjugates X
    assert np.abs(collect_density_cross(b, C1, C0, f) - X.conj()).max() < 1e-14 * scale
    # plane waves: component 0 at sphere point 2, component 1 at point 6, k 0
    a, c = 0.3 - 0.5j, 0.8 + 0.1j
    P0, P1 = np.zeros((K, 1, G), complex), np.zeros((K, 1, G), complex)
    P0[0, 0, 2], P1[0, 0, 6] = a, c
    fp = np.zeros((K, 1))
    fp[0, 0] = 0.75
    shape = np.array(b.shape)
    r = np.stack(np.meshgrid(*[np.arange(s) for s in b.shape], indexing="ij"), -1)
    phase = np.exp(2j * np.pi * ((b.iG[0, 6] - b.iG[0, 2]) * r / shape).sum(-1))
    want = 0.75 * b.w_sk[0] / b.volume * c * np.conj(a) * phase
    assert np.abs(collect_density_cross(b, P0, P1, fp) - want).max() < 1e-14


def test_harmonics_closed_forms_and_orthonormality():
    """Real solid harmonics: closed forms for l <= 2 (index l^2 + l + m,
    (y, z, x) at l = 1) and orthonormality over the unit sphere by
    Gauss-Legendre in cos(theta) times a uniform phi grid (exact for l <= 3)."""
    from oracle.projectors import real_solid_harmonics

    rng = np.random.default_rng(0)
    r = rng.standard_normal((50, 3))
    x, y, z = r[:, 0], r[:, 1], r[:, 2]
    r2 = (r**2).sum(-1)
    pi = np.pi
    want = [np.full_like(x, 0.5 / np.sqrt(pi)),
            np.sqrt(3 / (4 * pi)) * y, np.sqrt(3 / (4 * pi)) * z, np.sqrt(3 / (4 * pi)) * x,
            np.sqrt(15 / (4 * pi)) * x * y, np.sqrt(15 / (4 * pi)) * y * z,
            np.sqrt(5 / (16 * pi)) * (3 * z * z - r2), np.sqrt(15 / (4 * pi)) * x * z,
            np.sqrt(15 / (16 * pi)) * (x * x - y * y)]
    Y = real_solid_harmonics(2, r)
    assert np.abs(Y - np.array(want)).max() < 1e-13
    ct, wt = np.polynomial.legendre.leggauss(8)
    phi = np.arange(16) * 2 * np.pi / 16
    st = np.sqrt(1 - ct**2)
    pts = np.stack([st[:, None] * np.cos(phi), st[:, None] * np.sin(phi),
                    np.broadcast_to(ct[:, None], (8, 16))], -1).reshape(-1, 3)
    w = (wt[:, None] * np.full(16, 2 * np.pi / 16)).reshape(-1)
    Y = real_solid_harmonics(3, pts)
    assert np.abs((Y * w) @ Y.T - np.eye(16)).max() < 1e-13


def test_quintic_spline():
    """Blip coefficients: the spline reproduces its samples at the nodes,
    meets math/test_quintic_spline.py's bound |f - interp| < dx^4 on 2001
    points, is mirror-symmetric at 0 (c[-k] = c[k]) and has zero third and
    fourth derivatives at the last node."""
    from scipy.interpolate import BSpline

    from oracle.projectors import BLIP_SCALE, blip_eval, quintic_coeff

    dx = 0.1
    xs = np.arange(0.0, 40.0, dx)

    def f(x):
        return np.exp(-np.sin(0.01 * x * x)) * np.cos(0.1 * x)

    c = quintic_coeff(f(xs))
    assert c.shape == (len(xs) + 4,)
    assert np.abs(blip_eval(c, xs[:-1], dx) - f(xs[:-1])).max() < 1e-12
    fine = np.linspace(0.0, xs.max() - 1e-6 * dx, 2001)
    assert np.linalg.norm(blip_eval(c, fine, dx) - f(fine)) < dx**4
    assert c[0] == c[4] and c[1] == c[3]
    knots = (np.arange(len(c) + 6) - 5.0) * dx
    s = BSpline(knots, c * BLIP_SCALE, 5)
    scale = np.abs(s(xs, 2)).max()
    for d in (3, 4):
        assert abs(s(xs[-1], d)) < 1e-9 * scale / dx ** (d - 2)
    # batched samples keep their trailing shape
    cb = quintic_coeff(np.stack([f(xs), 2 * f(xs)]))
    assert cb.shape == (len(xs) + 4, 2) and np.allclose(cb[:, 1], 2 * c)


def test_spline_projectors_match_gaussian(small):
    """Projectors from sampled Gaussian radials (spacing 0.02, as the
    reference's RadialFunction.DG) equal the analytic Gaussian projectors of
    the synthetic cells to the spline's accuracy."""
    from oracle.hamiltonian import projectors as gauss_projectors
    from oracle.projectors import projectors, quintic_coeff

    spec, b = small
    dG = 0.02
    G = np.arange(int(np.ceil(np.sqrt(2 * spec.ke_cutoff) * 1.2 / dG)) + 5) * dG
    l_rf = [l for rc in spec.proj_rc for l in spec.proj_l]
    samples = np.stack([np.exp(-0.5 * rc**2 * G**2) for rc in spec.proj_rc for _ in spec.proj_l])
    coeff = quintic_coeff(samples)
    want = gauss_projectors(b)
    got = projectors(b, spec.positions, coeff, dG, l_rf, spec.volume)
    assert got.shape == want.shape
    assert np.abs(got - want).max() < 1e-8 * np.abs(want).max()


def test_tau_branch_closed_forms(small):
    """Oracle meta-GGA term -1/2 sum_dir grad(V_tau grad C): a constant
    V_tau = c gives c KE C; Hermitian for a real V_tau; <C|H_tau|C> >= 0 for
    V_tau >= 0."""
    from oracle.hamiltonian import apply_tau

    spec, b = small
    K, G = b.fft_index.shape
    rng = np.random.default_rng(9)
    C = rng.standard_normal((K, 3, G)) + 1j * rng.standard_normal((K, 3, G))
    b.constrain(C)
    got = apply_tau(b, np.full(b.shape, 0.7), C)
    want = 0.7 * b.KE[:, None, :] * C
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    V = 1.0 + 0.5 * rng.uniform(-1, 1, b.shape)  # positive
    A = rng.standard_normal((K, 2, G)) + 1j * rng.standard_normal((K, 2, G))
    b.constrain(A)
    HA, HC = apply_tau(b, V, A), apply_tau(b, V, C)
    lhs = np.einsum("kag,kbg->kab", A.conj(), HC)
    rhs = np.einsum("kbg,kag->kab", C.conj(), HA).conj()
    assert np.abs(lhs - rhs).max() < 1e-11 * np.abs(lhs).max()
    diag = np.einsum("kbg,kbg->kb", C.conj(), HC)
    assert np.all(diag.real > 0) and np.abs(diag.imag).max() < 1e-10 * diag.real.max()


def test_tau_branch_vs_dense(small):
    """Oracle apply_tau (through the FFT) against the dense brute force
    1/2 (G + k) . (G' + k) V~tau(G - G') on the in-sphere plane waves."""
    from oracle.dense import dense_hamiltonian
    from oracle.hamiltonian import apply_tau

    spec, b = small
    K, G = b.fft_index.shape
    rng = np.random.default_rng(10)
    V = rng.uniform(0.2, 1.0, b.shape)
    C = rng.standard_normal((K, 2, G)) + 1j * rng.standard_normal((K, 2, G))
    b.constrain(C)
    got = apply_tau(b, V, C)
    zero = np.zeros(b.shape)
    for ik in range(K):
        n = int(b.n[ik])
        H0 = dense_hamiltonian(b, ik, V=zero, D=np.zeros_like(spec.D))
        Ht = dense_hamiltonian(b, ik, V=zero, D=np.zeros_like(spec.D), V_tau=V)
        want = C[ik, :, :n] @ (Ht - H0).T
        assert np.abs(got[ik, :, :n] - want).max() < 1e-12 * np.abs(want).max()


def test_radial_transform_gaussian():
    """Oracle radial transform of exp(-r^2 / 2 s^2) (r^l removed) equals
    (2 pi)^(3/2) s^(2l + 3) exp(-s^2 G^2 / 2) for l = 0..3 (Simpson on an odd
    grid to 12 s)."""
    from oracle.projectors import radial_transform

    s = 0.8
    r = np.linspace(0.0, 12 * s, 2401)
    dr = np.full_like(r, r[1] - r[0])
    l_rf = [0, 1, 2, 3]
    f = np.stack([np.exp(-0.5 * r**2 / s**2)] * 4)
    ft = radial_transform(r, dr, f, l_rf, 6.0, 0.02)
    G = np.arange(ft.shape[1]) * 0.02
    for i, l in enumerate(l_rf):
        want = (2 * np.pi) ** 1.5 * s ** (2 * l + 3) * np.exp(-0.5 * s**2 * G**2)
        assert np.abs(ft[i] - want).max() < 1e-10 * want.max()
