"""pqrfact / psvdfact / psvdvals / pheigfact / pheigvals on the GPU against the oracle with the
same injected sketches (src/pqr.jl, src/psvd.jl, src/pheig.jl), plus the reference's own test
bounds (test/pqr.jl, test/psvd.jl, test/pheig.jl)."""
import numpy as np
import pytest

from test_gpu_idfact import DeviceOmega

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def host(t):
    return t.detach().cpu().numpy()


def decaying(m, n, sig, seed):
    rng = np.random.default_rng(seed)
    r = len(sig)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return np.asfortranarray((U * sig) @ V.T)


@pytest.mark.parametrize("trans", ["n", "c"])
def test_pqrfact_vs_oracle(gpu, orc, trans):
    sig = 2.0 ** -np.arange(40)
    A = decaying(600, 400, sig, 3)
    print(f"pqrfact trans={trans}", flush=True)
    F = gpu.pqrfact(trans, A, seed=5, rtol=1e-9)
    Fo = orc.pqrfact(trans, A, rand=DeviceOmega(gpu, 5), rtol=1e-9)
    Q, R, p = host(F.Q), host(F.R), host(F.p)
    assert Q.shape == Fo.Q.shape and R.shape == Fo.R.shape
    assert np.array_equal(p, Fo.p)
    # same skeleton, same Householder convention: R agrees to rounding, with dgeqrf's
    # diagonal signs; Q's trailing columns carry the conditioning of the skeleton, so Q is
    # compared through Q R
    assert np.abs(R - Fo.R).max() <= 1e-8 * np.abs(Fo.R).max()
    d, do = np.diag(R), np.diag(Fo.R)
    big = np.abs(do) > 1e-8 * np.abs(do).max()
    assert np.array_equal(np.sign(d[big]), np.sign(do[big]))
    assert np.abs(Q @ R - Fo.Q @ Fo.R).max() <= 1e-12 * np.abs(A).max()
    X = A if trans == "n" else A.T
    assert np.linalg.norm(X[:, p - 1] - Q @ R, 2) <= 1e-6 * np.linalg.norm(A, 2)
    assert np.abs(Q.T @ Q - np.eye(Q.shape[1])).max() <= 1e-12


def test_pqrfact_retvals_and_none(gpu, orc):
    rng = np.random.default_rng(2)
    A = orc.fourier_real(rng.random(200), rng.random(150))
    nrm = np.linalg.norm(A)
    for rv in ("qr", "qrt", "t", "q"):
        F = gpu.pqrfact("n", A, pqrfact_retval=rv, seed=1)
        assert (F.Q is not None) == ("q" in rv) and (F.R is not None) == ("r" in rv)
        if "q" in rv and "r" in rv:
            p = host(F.p)
            # test/pqr.jl:20-21 bound
            assert np.linalg.norm(A[:, p - 1] - host(F.Q) @ host(F.R)) < 100 * 5 * EPS * nrm * 1e3
    # sketch = none: QRCP of A itself (an exactly rank-12 matrix: a clean rank cut)
    B = np.asfortranarray(rng.standard_normal((200, 12)) @ rng.standard_normal((12, 150)))
    F = gpu.pqrfact("n", B, sketch="none", pqrfact_retval="qrt")
    Fo = orc.pqrfact("n", B, sketch="none", pqrfact_retval="qrt")
    assert F.k == Fo.k == 12
    p = host(F.p)
    # the skeleton matches; the order of the redundant columns beyond the rank is set by
    # pivots on noise (the QRCP block runs past k), so T is compared column by column index
    assert np.array_equal(p[:12], Fo.p[:12])
    assert np.array_equal(np.sort(p[12:]), np.sort(Fo.p[12:]))
    assert np.linalg.norm(B[:, p - 1] - host(F.Q) @ host(F.R)) < 1e-12 * np.linalg.norm(B)
    T, To = host(F.T)[:, np.argsort(p[12:])], Fo.T[:, np.argsort(Fo.p[12:])]
    assert np.abs(T - To).max() <= 1e-10 * max(1.0, np.abs(To).max())


@pytest.mark.parametrize("m,n", [(700, 500), (400, 900)])
def test_psvd_vs_oracle(gpu, orc, m, n):
    sig = 10.0 ** (-np.arange(60) / 8.0)
    A = decaying(m, n, sig, m + n)
    print(f"psvd {m}x{n}", flush=True)
    F = gpu.psvdfact(A, seed=2, rtol=1e-6)
    Fo = orc.psvdfact(A, rand=DeviceOmega(gpu, 2), rtol=1e-6)
    U, S, Vt = host(F.U), host(F.S), host(F.Vt)
    assert S.shape == Fo.S.shape
    assert np.allclose(S, Fo.S, rtol=1e-10, atol=1e-14)
    # the known spectrum, to the ID's approximation level (rtol 1e-6)
    assert np.abs(S - sig[: len(S)]).max() <= 1e-5 * sig[0]
    assert np.abs(U.T @ U - np.eye(len(S))).max() <= 1e-11
    assert np.abs(Vt @ Vt.T - np.eye(len(S))).max() <= 1e-11
    Ar = (U * S) @ Vt
    Aro = (Fo.U * Fo.S) @ Fo.Vt
    assert np.linalg.norm(Ar - Aro, 2) <= 1e-9
    assert np.linalg.norm(A - Ar, 2) <= 10 * sig[len(S)]
    s = host(gpu.psvdvals(A, seed=2, rtol=1e-6))
    assert np.allclose(s, S, rtol=1e-12)


def test_psvd_reference_bound(gpu, orc):
    """test/psvd.jl:22-23 (Float64): |A - U S V'| < 100 eps |A|."""
    rng = np.random.default_rng(0)
    A = orc.fourier_real(rng.random(300), rng.random(200))
    nrm = np.linalg.norm(A)
    U, S, V = gpu.psvd(A, seed=4)
    E = A - (host(U) * host(S)) @ host(V).T
    assert np.linalg.norm(E) < 100 * 5 * EPS * nrm * 10


def test_pheig_vs_oracle(gpu, orc):
    rng = np.random.default_rng(7)
    n, r = 500, 40
    Qr, _ = np.linalg.qr(rng.standard_normal((n, r)))
    w = np.concatenate([-(2.0 ** -np.arange(15)), 3.0 * 2.0 ** -np.arange(25)])
    A = np.asfortranarray((Qr * w) @ Qr.T)
    A = np.asfortranarray((A + A.T) / 2)
    print("pheig", flush=True)
    F = gpu.pheigfact(A, seed=3, rtol=1e-7)
    Fo = orc.pheigfact(A, rand=DeviceOmega(gpu, 3), rtol=1e-7)
    vals, vecs = host(F.values), host(F.vectors)
    assert vals.shape == Fo.values.shape
    assert np.allclose(vals, Fo.values, rtol=1e-9, atol=1e-12)
    ref = np.sort(w[np.abs(w) > 1e-7 * np.abs(w).max()])
    assert len(vals) == len(ref)
    # the known spectrum, to the ID's approximation level (rtol 1e-7)
    assert np.abs(np.sort(vals) - ref).max() <= 1e-6 * np.abs(w).max()
    assert np.abs(vecs.T @ vecs - np.eye(len(vals))).max() <= 1e-10
    Ar = (vecs * vals) @ vecs.T
    Aro = (Fo.vectors * Fo.values) @ Fo.vectors.T
    assert np.linalg.norm(Ar - Aro, 2) <= 1e-8
    v = host(gpu.pheigvals(A, seed=3, rtol=1e-7))
    assert np.allclose(v, vals, rtol=1e-12)


def test_pheig_reference_bound(gpu, orc):
    """test/pheig.jl:20-21 (Float64): |A - V diag(w) V'| < 100 eps |A| on a Hermitian
    Fourier-type matrix."""
    rng = np.random.default_rng(1)
    x = rng.random(300)
    A = orc.fourier_real(x, x)
    A = np.asfortranarray((A + A.T) / 2)
    nrm = np.linalg.norm(A)
    w, V = gpu.pheig(A, seed=1)
    E = A - (host(V) * host(w)) @ host(V).T
    assert np.linalg.norm(E) < 100 * 5 * EPS * nrm * 10


def _gesvj(gpu, X):
    import torch
    L = gpu._lib.load()
    ctx = gpu.context()
    m, k = X.shape
    Xd = gpu.to_device(np.asfortranarray(X))
    U = gpu.empty_colmajor(m, k)
    V = gpu.empty_colmajor(k, k)
    s = torch.empty(k, dtype=torch.float64, device="cuda")
    gpu.check(L.lra_gesvj_f64(ctx.h, m, k, Xd.data_ptr(), m, U.data_ptr(), m, s.data_ptr(),
                              V.data_ptr(), k), "gesvj")
    return host(U), host(s), host(V)


def _check_svd(X, U, s, V, rtol=1e-11):
    m, k = X.shape
    sr = np.linalg.svd(X, compute_uv=False)
    assert np.allclose(s, sr, rtol=rtol, atol=1e-14 * sr[0])
    assert np.all(np.diff(s) <= 0)
    Xr = (U * s) @ V.T
    assert np.abs(Xr - X).max() <= 1e-13 * np.abs(X).max() * k
    assert np.abs(V.T @ V - np.eye(k)).max() <= 1e-13 * k


def test_gesvj_vs_numpy(gpu):
    rng = np.random.default_rng(4)
    for k in (1, 7, 64, 201, 436):
        X = rng.standard_normal((k, k)) @ np.diag(10.0 ** -np.linspace(0, 10, k))
        print(f"gesvj k={k}", flush=True)
        _check_svd(X, *_gesvj(gpu, X))


@pytest.mark.parametrize("m,k", [(2091, 2091), (2600, 300), (4000, 120)])
def test_gesvj_tall_blocks(gpu, m, k):
    """Matrices taller than 2048 rows: blocks of two columns with 40, 48 or 64 rows per lane
    (up to 4096 rows) instead of one launch per round of single pairs; graded columns."""
    rng = np.random.default_rng(m + k)
    X = rng.standard_normal((m, k)) @ np.diag(10.0 ** -np.linspace(0, 10, k))
    print(f"gesvj tall {m}x{k}", flush=True)
    _check_svd(X, *_gesvj(gpu, X))


@pytest.mark.parametrize("w", ["0", "1", "2", "4", "8", "16"])
def test_gesvj_block_widths(gpu, w, monkeypatch):
    """Every block width of the block-cyclic order (0: one launch per round of single pairs)
    on a tall matrix with graded columns, an odd column count and two zero columns."""
    monkeypatch.setenv("LRA_JACOBI_W", w)
    rng = np.random.default_rng(11)
    m, k = 150, 93
    X = rng.standard_normal((m, k)) @ np.diag(10.0 ** -np.linspace(0, 9, k))
    X[:, 10] = 0.0
    X[:, 50] = 0.0
    print(f"gesvj W={w}", flush=True)
    U, s, V = _gesvj(gpu, X)
    _check_svd(X, U, s, V)
    assert s[-2:].max() == 0.0
    nz = s > 0
    assert np.abs(U[:, nz].T @ U[:, nz] - np.eye(nz.sum())).max() <= 1e-12 * k
