"""Factor application (src/id.jl:94-331, src/pqr.jl:86-175, src/psvd.jl:61-124,
src/pheig.jl:54-100) against dense numpy products of the same factors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def test_idpackedv_and_id_apply(gpu, orc):
    rng = np.random.default_rng(0)
    A = orc.fourier_real(rng.random(180), rng.random(140))
    V = gpu.idfact(A, seed=1)
    sk, rd, T = host(V.sk), host(V.rd), host(V.T)
    k, n = V.k, V.n
    Vd = np.zeros((k, n))
    Vd[:, sk - 1] = np.eye(k)
    Vd[:, rd - 1] = T
    assert np.allclose(host(V.matrix()), Vd, atol=1e-14)
    X = rng.standard_normal((n, 5))
    assert np.allclose(host(V.apply(X)), Vd @ X, atol=1e-12)
    Yk = rng.standard_normal((k, 3))
    assert np.allclose(host(V.apply(Yk, "c")), Vd.T @ Yk, atol=1e-12)
    F = gpu.ID.from_matrix(A, V)
    Cd = A[:, sk - 1]
    assert np.allclose(host(F.apply(X)), Cd @ (Vd @ X), atol=1e-11)
    Z = rng.standard_normal((180, 4))
    assert np.allclose(host(F.apply(Z, "c")), Vd.T @ (Cd.T @ Z), atol=1e-11)
    z = Z[:, 0].copy()  # a vector input gives a vector
    assert host(F.apply(z, "c")).shape == (n,)
    assert np.allclose(host(F.apply(z, "c")), Vd.T @ (Cd.T @ z), atol=1e-11)
    assert np.allclose(host(F.apply(X[:, 0].copy())), Cd @ (Vd @ X[:, 0]), atol=1e-11)
    assert np.allclose(host(V.apply(X[:, 1].copy())), Vd @ X[:, 1], atol=1e-12)


def test_partial_factor_apply(gpu, orc):
    rng = np.random.default_rng(1)
    A = orc.fourier_real(rng.random(200), rng.random(150))
    X = rng.standard_normal((150, 4))
    Fq = gpu.pqrfact("n", A, seed=2)
    Q, R, p = host(Fq.Q), host(Fq.R), host(Fq.p)
    M = np.empty_like(A)
    M[:, p - 1] = Q @ R
    assert np.allclose(host(Fq.apply(X)), M @ X, atol=1e-11)
    Fs = gpu.psvdfact(A, seed=3)
    Md = (host(Fs.U) * host(Fs.S)) @ host(Fs.Vt)
    assert np.allclose(host(Fs.apply(X)), Md @ X, atol=1e-11)
    Y = rng.standard_normal((200, 2))
    assert np.allclose(host(Fs.apply(Y, "c")), Md.T @ Y, atol=1e-11)
    S = A[:150, :] @ A[:150, :].T
    S = (S + S.T) / 2  # exactly symmetric
    Fe = gpu.pheigfact(S, seed=4)
    Me = (host(Fe.vectors) * host(Fe.values)) @ host(Fe.vectors).T
    Xe = rng.standard_normal((150, 3))
    assert np.allclose(host(Fe.apply(Xe)), Me @ Xe, atol=1e-9 * np.abs(S).max())


@pytest.mark.parametrize("trans", ["n", "c"])
def test_partial_qr_arithmetic(gpu, trans):
    """PartialQR products on both sides, with and without adjoint, Matrix(A) and the
    least-squares solve (src/pqr.jl:56, 86-175) against numpy on the same factors."""
    rng = np.random.default_rng(5)
    m, n = 260, 190
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = np.asfortranarray((U * np.exp(-0.1 * np.arange(n))) @ V.T)
    print(f"partial qr arithmetic trans={trans}", flush=True)
    F = gpu.pqrfact(trans, A, rtol=1e-6, seed=7)
    Q, R, p = host(F.Q), host(F.R), host(F.p)
    P, N = Q.shape[0], R.shape[1]
    assert not np.tril(R, -1).any()  # upper trapezoidal, stored dense
    M = np.empty((P, N))
    M[:, p - 1] = Q @ R
    assert np.allclose(host(F.matrix()), M, atol=1e-13)
    X = rng.standard_normal((N, 4))
    Y = rng.standard_normal((P, 3))
    Z = rng.standard_normal((5, P))
    W = rng.standard_normal((2, N))
    tol = 1e-12 * np.abs(M).max() * max(P, N)
    assert np.abs(host(F.apply(X)) - M @ X).max() <= tol
    assert np.abs(host(F.apply(Y, "c")) - M.T @ Y).max() <= tol
    assert np.abs(host(F.rapply(Z)) - Z @ M).max() <= tol
    assert np.abs(host(F.rapply(W, "c")) - W @ M.T).max() <= tol
    # A \ B = P R^+ Q' B, the minimum-norm least-squares solution: against numpy's
    # least-squares solve of the same R, and the normal equations M'(M x - B) = 0
    B = rng.standard_normal((P, 3))
    x = host(F.ldiv(B))
    xr = np.empty((N, 3))
    xr[p - 1, :] = np.linalg.lstsq(R, Q.T @ B, rcond=None)[0]
    assert np.abs(x - xr).max() <= 1e-8 * np.abs(xr).max(), np.abs(x - xr).max()
    assert np.abs(M.T @ (M @ x - B)).max() <= 1e-8 * np.abs(M).max() * np.abs(B).max()
    b = rng.standard_normal(P)
    assert host(F.ldiv(b)).ndim == 1
    assert np.allclose(host(F.ldiv(b)), np.linalg.lstsq(M, b, rcond=1e-9)[0],
                       atol=1e-8 * np.abs(xr).max())


def test_psvd_pheig_cur_arithmetic(gpu, orc):
    """Right products, Matrix(A) and the pseudo-inverse solve of PartialSVD and
    PartialHermEigen (src/psvd.jl:38, 99-124; src/pheig.jl:31, 114-152) and CUR's products
    (src/cur.jl:153, 436-523) against numpy on the same factors."""
    rng = np.random.default_rng(6)
    A = orc.fourier_real(rng.random(210), rng.random(160))
    m, n = A.shape
    print("psvd/pheig/cur arithmetic", flush=True)
    Fs = gpu.psvdfact(A, rtol=1e-8, seed=3)
    U, S, Vt = host(Fs.U), host(Fs.S), host(Fs.Vt)
    Md = (U * S) @ Vt
    assert np.allclose(host(Fs.matrix()), Md, atol=1e-13)
    Z = rng.standard_normal((3, m))
    W = rng.standard_normal((2, n))
    assert np.allclose(host(Fs.rapply(Z)), Z @ Md, atol=1e-11)
    assert np.allclose(host(Fs.rapply(W, "c")), W @ Md.T, atol=1e-11)
    B = rng.standard_normal((m, 3))
    x = host(Fs.ldiv(B))
    assert np.allclose(x, Vt.T @ ((U.T @ B) / S[:, None]), rtol=1e-10,
                       atol=1e-12 * np.abs(x).max())
    bv = rng.standard_normal(m)  # vectors act as one-column matrices
    assert host(Fs.apply(rng.standard_normal(n))).shape == (m,)
    assert np.allclose(host(Fs.ldiv(bv)), Vt.T @ ((U.T @ bv) / S), rtol=1e-10,
                       atol=1e-12 * np.abs(x).max())
    # the minimum-norm least-squares solution of the factored matrix
    assert np.allclose(x, np.linalg.lstsq(Md, B, rcond=S[-1] / S[0] / 2)[0],
                       atol=1e-6 * np.abs(x).max())

    Sm = A[:150, :] @ A[:150, :].T
    Sm = (Sm + Sm.T) / 2
    Fe = gpu.pheigfact(Sm, rtol=1e-8, seed=4)
    w, Ve = host(Fe.values), host(Fe.vectors)
    Me = (Ve * w) @ Ve.T
    sc = np.abs(Sm).max()
    assert np.allclose(host(Fe.matrix()), Me, atol=1e-13 * sc)
    Ze = rng.standard_normal((4, 150))
    assert np.allclose(host(Fe.rapply(Ze)), Ze @ Me, atol=1e-11 * sc)
    Be = rng.standard_normal((150, 2))
    ve = rng.standard_normal(150)
    assert np.allclose(host(Fe.apply(ve)), Me @ ve, atol=1e-11 * sc)
    xe = host(Fe.ldiv(Be))
    assert np.allclose(xe, Ve @ ((Ve.T @ Be) / w[:, None]), rtol=1e-10,
                       atol=1e-12 * np.abs(xe).max())

    F = gpu.CUR.from_packed(A, gpu.curfact(A, rtol=1e-6, seed=5))
    Cm, Rm = host(F.C), host(F.R)
    Um = host(F.U.matrix())
    Mc = Cm @ Um @ Rm
    # rounding scale of C U R products: |C| |U| |R| k^2 eps
    tol = 1e-14 * np.abs(Cm).max() * np.abs(Um).max() * np.abs(Rm).max() * Um.shape[0] ** 2
    assert np.abs(host(F.matrix()) - Mc).max() <= tol
    assert np.abs(host(F.rapply(Z)) - Z @ Mc).max() <= tol * np.abs(Z).sum(1).max()
    assert np.abs(host(F.rapply(W, "c")) - W @ Mc.T).max() <= tol * np.abs(W).sum(1).max()


def test_id_right_products(gpu, orc):
    """The right products of IDPackedV and ID (src/id.jl:154-200, 302-330), i.e. the
    B*V, B*V', B*A and B*A' identities of test/id.jl:34-57, and Matrix(V), Matrix(A), against
    numpy on the same factors."""
    rng = np.random.default_rng(11)
    A = orc.fourier_real(rng.random(170), rng.random(130))
    m, n = A.shape
    V = gpu.idfact(A, seed=2, rtol=1e-10)
    sk, rd, T = host(V.sk), host(V.rd), host(V.T)
    k = V.k
    Vd = np.zeros((k, n))
    Vd[:, sk - 1] = np.eye(k)
    Vd[:, rd - 1] = T
    assert np.array_equal(host(V.matrix()), Vd)
    Bk = rng.standard_normal((5, k))
    Bn = rng.standard_normal((4, n))
    assert np.allclose(host(V.rapply(Bk)), Bk @ Vd, atol=1e-12)        # B * V
    assert np.allclose(host(V.rapply(Bn, "c")), Bn @ Vd.T, atol=1e-12)  # B * V'
    F = gpu.ID.from_matrix(A, V)
    Ad = A[:, sk - 1] @ Vd
    assert np.allclose(host(F.matrix()), Ad, atol=1e-12)
    Bm = rng.standard_normal((3, m))
    assert np.allclose(host(F.rapply(Bm)), Bm @ Ad, atol=1e-11)        # B * A
    assert np.allclose(host(F.rapply(Bn, "c")), Bn @ Ad.T, atol=1e-11)  # B * A'
    x = rng.standard_normal(m)
    y = host(F.rapply(x))
    assert y.shape == (n,) and np.allclose(y, x @ Ad, atol=1e-11)
    # the reference's bound on the approximation (test/id.jl:34-57 with rtol = 1e-10)
    assert np.linalg.norm(A - Ad, 2) <= 1e-8 * np.linalg.norm(A, 2)
