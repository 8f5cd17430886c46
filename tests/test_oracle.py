"""CPU pins of the parity oracle (oracle/).  Every oracle function is checked against an
independent answer: OpenBLAS's own dlaqps (the routine the reference ccall's), scipy's dgeqp3,
the reference README's published ranks, numpy.fft for the SRFT, dense constructions for the
subset / sparse-Gaussian sketches, numpy.linalg for norms, SVD and eig, and the residual bounds
of the reference's own tests (test/id.jl, test/pqr.jl, test/psvd.jl, test/pheig.jl,
test/cur.jl, test/snorm.jl)."""
import numpy as np
import pytest
import scipy.linalg as sla

EPS = np.finfo(np.float64).eps


def graded(m, n, seed, decay=0.8):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((m, min(m, n)))
    V = rng.standard_normal((min(m, n), n))
    return np.asfortranarray((U * decay ** np.arange(min(m, n))) @ V)


# ---------------------------------------------------------------------------------------
# dlaqps restatement vs OpenBLAS dlaqps (src/lapack.jl:129-136 calls the latter)

def _cancelling(m, n, seed):
    """Columns nearly parallel to the first ones: norm downdates cancel and force the
    recompute path (dlaqps' lsticc, early block end)."""
    rng = np.random.default_rng(seed)
    base = rng.standard_normal((m, 4))
    A = np.empty((m, n))
    for j in range(n):
        A[:, j] = base[:, j % 4] * (1 + j / n) + 1e-9 * rng.standard_normal(m)
    return np.asfortranarray(A)


@pytest.mark.parametrize("maker,m,n,offset,nb", [
    (lambda: graded(40, 200, 1), 40, 200, 0, 32),
    (lambda: graded(60, 300, 2), 60, 300, 10, 16),
    (lambda: graded(200, 120, 3, 0.9), 200, 120, 7, 32),
    (lambda: _cancelling(50, 80, 4), 50, 80, 0, 32),
    (lambda: np.asfortranarray(np.random.default_rng(5).standard_normal((30, 30))), 30, 30, 3, 8),
])
def test_laqps_matches_openblas(orc, maker, m, n, offset, nb):
    A = maker()
    norms = np.linalg.norm(A[offset:, :], axis=0)
    jp = np.arange(1, n + 1, dtype=np.int64)
    r1 = orc.laqps(A, offset, nb, jp, norms, norms)
    r2 = orc.laqps(A, offset, nb, jp, norms, norms, use_openblas=True)
    assert r1[5] == r2[5]                       # kb
    assert np.array_equal(r1[1], r2[1])         # jpvt
    assert np.abs(r1[0] - r2[0]).max() <= 1e-12 * np.abs(A).max()
    assert np.allclose(r1[2], r2[2], rtol=1e-12, atol=1e-14)
    kb = r1[5]
    assert np.allclose(r1[3][kb:], r2[3][kb:], rtol=1e-10)


def test_laqps_recompute_path_exercised(orc):
    A = _cancelling(50, 80, 4)
    norms = np.linalg.norm(A, axis=0)
    _, _, _, _, _, kb = orc.laqps(A, 0, 32, np.arange(1, 81, dtype=np.int64), norms, norms)
    assert kb < 32  # the block ended early at a flagged downdate


@pytest.mark.parametrize("m,n", [(40, 300), (120, 80), (64, 64)])
def test_geqp3_adap_full_rank_matches_dgeqp3(orc, m, n):
    B = np.asfortranarray(np.random.default_rng(m * n).standard_normal((m, n)))
    Bf, p, tau, k = orc.geqp3_adap(B, -1, 0.0, 0.0, 32)
    Q, R, P = sla.qr(B, pivoting=True, mode="economic")
    assert k == min(m, n)
    assert np.array_equal(p - 1, P)
    assert np.allclose(np.abs(np.diag(Bf[:k, :k])), np.abs(np.diag(R)), rtol=1e-10)


def test_geqp3_adap_openblas_and_restated_agree_on_rank(orc):
    for seed in range(4):
        B = graded(72, 2000, seed, 0.85)
        a = orc.geqp3_adap(B, -1, 0.0, 1e-10, 32)
        b = orc.geqp3_adap(B, -1, 0.0, 1e-10, 32, use_openblas=True)
        assert a[3] == b[3] and np.array_equal(a[1], b[1])


def test_geqp3_adap_special_cases(orc):
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(30), rng.standard_normal(200)
    _, p, _, k = orc.geqp3_adap(np.outer(x, y), -1, 0.0, 1e-12, 32)
    assert k == 1 and p[0] == np.argmax(np.abs(y)) + 1          # rank one
    L = rng.standard_normal((40, 7)) @ rng.standard_normal((7, 300))
    _, _, _, k = orc.geqp3_adap(L, -1, 0.0, 1e-12, 32)
    assert k == 7                                                 # exactly rank 7
    _, _, _, k = orc.geqp3_adap(graded(60, 400, 1), 20, 0.0, 0.0, 32)
    assert k == 20                                                # rank cap


# ---------------------------------------------------------------------------------------
# Hilbert ranks published in the reference README (README.md:107,122,125)

def test_hilbert_ranks_match_readme(orc):
    A = orc.hilbert(1024)
    ks = [orc.idfact("n", A, rand=orc.Rand(s)).k for s in range(5)]
    assert set(ks) <= {26, 27}
    ks = [orc.idfact("n", A, rand=orc.Rand(s), rtol=1e-12).k for s in range(3)]
    assert all(abs(k - 22) <= 1 for k in ks)
    assert orc.idfact("n", A, rand=orc.Rand(0), rank=20).k == 20


# ---------------------------------------------------------------------------------------
# T = R11 \ R12 and strong-RRQR swaps

def test_maxdet_t_matches_solve_triangular(orc):
    rng = np.random.default_rng(3)
    k, n = 30, 200
    R = np.triu(rng.standard_normal((k, n)))
    R[np.arange(k), np.arange(k)] += 4 * np.sign(R[np.arange(k), np.arange(k)])
    T = orc.maxdet_t(R)
    assert np.allclose(T, sla.solve_triangular(R[:, :k], R[:, k:]), rtol=1e-12, atol=1e-12)


def test_maxdet_t_blas_matches_restated(orc):
    """The CPU baseline's OpenBLAS dtrsm T equals the C restatement and scipy's solve."""
    rng = np.random.default_rng(4)
    k, n = 45, 310
    R = np.triu(rng.standard_normal((k, n)))
    R[np.arange(k), np.arange(k)] += 4 * np.sign(R[np.arange(k), np.arange(k)])
    T = orc.maxdet_t_blas(R)
    assert T.shape == (k, n - k)
    assert np.allclose(T, orc.maxdet_t(R), rtol=1e-12, atol=1e-12)
    assert np.allclose(T, sla.solve_triangular(R[:, :k], R[:, k:]), rtol=1e-12, atol=1e-12)
    # idfact with the OpenBLAS legs gives the restated idfact's factorization
    A = orc.hilbert(300)
    V1 = orc.idfact("n", A, rand=orc.Rand(5), rtol=1e-12)
    V2 = orc.idfact("n", A, rand=orc.Rand(5), rtol=1e-12, use_openblas=True)
    assert V1.k == V2.k and np.array_equal(V1.sk[:10], V2.sk[:10])


def test_maxdet_swaps_bound_T_and_keep_the_id(orc):
    rng = np.random.default_rng(11)
    A = rng.standard_normal((80, 12)) @ rng.standard_normal((12, 150))
    V = orc.idfact("n", A, rand=orc.Rand(2), rank=12)
    p = np.concatenate([V.sk, V.rd])
    # inflate T by swapping a skeleton column with a redundant one on purpose
    p2, T2, _, it = orc.maxdet_swapcols(p, V.T, 0.0, -1)
    assert np.abs(T2).max() <= 1.0 + 1e-12
    sk, rd = p2[:12], p2[12:]
    assert np.abs(A[:, rd - 1] - A[:, sk - 1] @ T2).max() <= 1e-9 * np.abs(A).max()


# ---------------------------------------------------------------------------------------
# structured sketches

def test_srft_l_rule(orc):
    assert orc.srft_l(65536, 310) == 256
    assert orc.srft_l(1000, 40) == 40
    assert orc.srft_l(97, 40) == 1
    assert orc.srft_l(12, 5) == 4


@pytest.mark.parametrize("m,n,k", [(12, 3, 5), (64, 5, 16), (60, 4, 9), (97, 3, 6)])
def test_srft_closed_form(orc, m, n, k):
    rng = np.random.default_rng(m + k)
    A = np.asfortranarray(rng.standard_normal((m, n)))
    d = np.where(rng.random(m) > 0.5, 1.0, -1.0)
    idx = rng.integers(1, m + 1, size=k)
    B = orc.srft_left(A, k, d, idx)
    F = np.fft.fft(d[:, None] * A, axis=0)   # DFT_m(d .* A[:, j]), no scaling
    i = 0
    while i < k:
        f = idx[i] - 1
        assert np.allclose(B[i], F[f].real, atol=1e-9 * np.sqrt(m))
        if i == k - 1:
            break
        assert np.allclose(B[i + 1], F[f].imag, atol=1e-9 * np.sqrt(m))
        i += 2


def test_subset_and_sprn(orc):
    rng = np.random.default_rng(1)
    m, n, k = 50, 7, 6
    A = np.asfortranarray(rng.standard_normal((m, n)))
    r = rng.integers(1, m + 1, size=k)
    assert np.array_equal(orc.subset_left(A, k, r), A[r - 1, :])
    perm = rng.permutation(m) + 1
    s = rng.standard_normal(m)
    S = np.zeros((k, m))
    idx = 0
    for i in range(1, k + 1):
        p = (m - i) // k + 1
        S[i - 1, perm[idx:idx + p] - 1] = s[idx:idx + p]
        idx += p
    assert idx == m
    assert np.allclose(orc.sprn_left(A, k, perm, s), S @ A, atol=1e-13)


# ---------------------------------------------------------------------------------------
# snorm (test/snorm.jl:12-25) and the factorizations' residual bounds (test/*.jl)

def test_snorm_matches_opnorm(orc):
    rng = np.random.default_rng(0)
    rtol = 100 * EPS
    A = rng.random((20, 10))
    nrm = np.linalg.norm(A, 2)
    assert abs(orc.snorm(A) - nrm) <= rtol * nrm
    A = rng.random((10, 10)); A = A + A.T
    nrm = np.linalg.norm(A, 2)
    assert abs(orc.snorm(A) - nrm) <= rtol * nrm
    B = rng.random((10, 10))
    nrm = np.linalg.norm(A - B, 2)
    assert abs(orc.snormdiff(A, B) - nrm) <= rtol * nrm


@pytest.fixture(scope="module")
def fourier(orc):
    rng = np.random.default_rng(0)
    return orc.fourier_real(rng.random(128), rng.random(64))


@pytest.mark.parametrize("sk", ["none", "randn", "sub", "srft", "sprn"])
def test_idfact_residual_bound(orc, fourier, sk):
    opts = orc.LRAOptions(maxdet_tol=0.0, sketch_randn_niter=1, sketch=sk)
    A = fourier
    nrm = np.linalg.norm(A)
    for trans in ("n", "c"):
        V = orc.idfact(trans, A, opts, rand=orc.Rand(4))
        E = orc.id_residual(A, V, trans)
        assert np.linalg.norm(E) < 100 * 5 * EPS * nrm      # test/id.jl:20-32
        assert np.abs(V.T).max() <= 1.0 + 1e-12              # maxdet_tol = 0


@pytest.mark.parametrize("sk", ["none", "randn", "srft"])
def test_pqr_psvd_pheig_cur_residuals(orc, fourier, sk):
    A = fourier
    opts = orc.LRAOptions(maxdet_tol=0.0, sketch_randn_niter=1, sketch=sk)
    tol = 100 * 5 * EPS * np.linalg.norm(A)
    F = orc.pqrfact("n", A, opts, rand=orc.Rand(1))
    assert np.linalg.norm(A - F.matrix()) < tol                  # test/pqr.jl:29
    S = orc.psvdfact(A, opts, rand=orc.Rand(1))
    assert np.linalg.norm(A - S.matrix()) < tol                  # test/psvd.jl:29
    s = orc.psvdvals(A, opts, rand=orc.Rand(1), rank=len(S.S), rtol=0.0)
    assert np.linalg.norm(s - S.S) < 100 * 5 * EPS * np.linalg.norm(s)
    H = orc.fourier_real(np.random.default_rng(2).random(128), np.random.default_rng(2).random(128))
    H = H + H.T
    E = orc.pheigfact(H, opts, rand=orc.Rand(1))
    assert np.linalg.norm(H - E.matrix()) < 100 * 5 * EPS * np.linalg.norm(H)   # test/pheig.jl:26
    rows, cols = orc.curfact(A, opts, rand=orc.Rand(1))
    # curfact returns skeletons only (CURPackedU): they must span A's column / row space to
    # the test/cur.jl:23 tolerance
    C, R = A[:, cols - 1], A[rows - 1, :]
    Qc, _ = np.linalg.qr(C)
    Qr, _ = np.linalg.qr(R.T)
    assert np.linalg.norm(A - Qc @ (Qc.T @ A)) < 1000 * 5 * EPS * np.linalg.norm(A)
    assert np.linalg.norm(A - (A @ Qr) @ Qr.T) < 1000 * 5 * EPS * np.linalg.norm(A)
    assert len(rows) == len(cols)


def test_psvd_recovers_known_singular_values(orc):
    rng = np.random.default_rng(7)
    m, n, r = 600, 400, 60
    sig = 10.0 ** (-8 * np.arange(r) / (r - 1))
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    A = (U * sig) @ V.T
    S = orc.psvdfact(A, rand=orc.Rand(3), rtol=1e-10)
    assert len(S.S) == r
    assert np.abs(S.S - sig).max() <= 1e-12


@pytest.mark.parametrize("trans", ["n", "c"])
@pytest.mark.parametrize("sketch", ["randn", "none", "sub"])
def test_oracle_prange_exact_rank(trans, sketch):
    """prange of an exactly rank-r matrix: k = r and Q spans the true range (the projector
    equals the one of the r leading left singular vectors of op(A))."""
    import oracle as orc
    rng = np.random.default_rng(3)
    m, n, r = 120, 90, 17
    A = rng.standard_normal((m, r)) @ rng.standard_normal((r, n))
    X = A if trans == "n" else A.T
    Q = orc.prange(trans, A, rtol=1e-12, sketch=sketch)
    assert Q.shape == (X.shape[0], r)
    U = np.linalg.svd(X)[0][:, :r]
    assert np.abs(Q @ Q.T - U @ U.T).max() <= 1e-10
    assert np.abs(Q.T @ Q - np.eye(r)).max() <= 1e-13


def test_oracle_prange_both_spans_both():
    import oracle as orc
    rng = np.random.default_rng(4)
    n, r = 80, 9
    A = rng.standard_normal((n, r)) @ rng.standard_normal((r, n))
    Q = orc.prange("b", A, rtol=1e-12)
    # range(A) + range(A') has dimension <= 2r; Q spans both
    assert Q.shape[1] <= 2 * r
    for X in (A, A.T):
        assert np.linalg.norm(X - Q @ (Q.T @ X), 2) <= 1e-10 * np.linalg.norm(X, 2)


def test_oracle_cur_build_exact_rank():
    """A = C W^+ R when rank(W) = rank(A) = k: the CUR factors of an exactly rank-k matrix
    from its skeleton reproduce it, and U is the pseudo-inverse of the intersection."""
    import oracle as orc
    rng = np.random.default_rng(5)
    m, n, k = 90, 70, 12
    A = rng.standard_normal((m, k)) @ rng.standard_normal((k, n))
    rows, cols = orc.curfact(A, rtol=1e-12)
    assert len(rows) == k and len(cols) == k
    C, (V, sinv, Ut), R = orc.cur_build(A, rows, cols)
    Wp = V @ np.diag(sinv) @ Ut
    W = A[np.ix_(np.asarray(rows) - 1, np.asarray(cols) - 1)]
    assert np.allclose(Wp, np.linalg.pinv(W), rtol=1e-8, atol=1e-10 * np.abs(Wp).max())
    assert np.abs(C @ Wp @ R - A).max() <= 1e-9 * np.abs(A).max()


# ---------------------------------------------------------------------------------------
# complex element type: zlaqps loop, complex sketches, complex idfact

def _zlowrank(m, n, r, seed, decay=None):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))
    V = rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r))
    s = np.ones(r) if decay is None else decay ** np.arange(r)
    return np.asfortranarray((U * s) @ V.conj().T)


def test_zgeqp3_adap_full_rank_matches_zgeqp3(orc):
    """The restated loop over OpenBLAS zlaqps against LAPACK zgeqp3 (scipy): same pivots and
    |R| on a full-rank complex matrix, and B[:, p] = Q R."""
    rng = np.random.default_rng(7)
    m, n = 60, 40
    B = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))) * np.exp(
        -0.15 * np.arange(n))
    Bf, p, tau, k = orc.zgeqp3_adap(B, -1, 0.0, 0.0, 8)
    assert k == n
    Q, R, P = sla.qr(B, pivoting=True, mode="economic")
    assert np.array_equal(p - 1, P)
    assert np.allclose(np.abs(np.diag(Bf)), np.abs(np.diag(R)), rtol=1e-10)
    Rt = np.triu(Bf[:n, :])
    Qo = np.linalg.qr(B[:, p - 1])[0]
    assert np.allclose(np.abs(Qo.conj().T @ B[:, p - 1]), np.abs(Rt), atol=1e-10)


def test_zgeqp3_adap_exact_rank_and_rank_cap(orc):
    B = _zlowrank(80, 70, 13, 3)
    _, p, _, k = orc.zgeqp3_adap(B, -1, 0.0, 1e-12, 32)
    assert k == 13
    _, p, _, k = orc.zgeqp3_adap(B, 5, 0.0, 1e-12, 32)
    assert k == 5


def test_zsketch_phi_closed_forms(orc):
    rng = np.random.default_rng(5)
    m, l = 96, 11
    x = rng.standard_normal((m, 3)) + 1j * rng.standard_normal((m, 3))
    g = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    d = g / np.abs(g)
    idx = rng.integers(1, m + 1, size=l)
    Phi = orc.zsketch_phi("srft", m, l, (d, idx))
    F = np.fft.fft(d[:, None] * x, axis=0)
    assert np.allclose(Phi @ x, F[idx - 1], atol=1e-12 * np.sqrt(m))
    r = rng.integers(1, m + 1, size=l)
    assert np.array_equal(orc.zsketch_phi("sub", m, l, r) @ x, x[r - 1])
    perm = rng.permutation(m) + 1
    sc = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    Ps = orc.zsketch_phi("sprn", m, l, (perm, sc))
    assert np.count_nonzero(Ps) == m and np.array_equal(np.sort(np.nonzero(Ps)[1]), np.arange(m))
    assert [np.count_nonzero(Ps[i]) for i in range(l)] == [(m - i - 1) // l + 1 for i in range(l)]


class _ZRand:
    def __init__(self, seed):
        self.rng = np.random.default_rng(seed)

    def phi(self, sketch, rnd, m, order):
        r = self.rng
        if sketch == "randn":
            return r.standard_normal((order, m)) + 1j * r.standard_normal((order, m))
        if sketch == "srft":
            g = r.standard_normal(m) + 1j * r.standard_normal(m)
            return _orc().zsketch_phi("srft", m, order, (g / np.abs(g), r.integers(1, m + 1, order)))
        if sketch == "sub":
            return _orc().zsketch_phi("sub", m, order, r.integers(1, m + 1, order))
        return _orc().zsketch_phi("sprn", m, order, (r.permutation(m) + 1,
                                                     r.standard_normal(m) + 1j * r.standard_normal(m)))


def _orc():
    import oracle
    return oracle


@pytest.mark.parametrize("sk", ["randn", "srft", "sub", "sprn", "none"])
def test_zidfact_residual_and_rank(orc, sk):
    """test/id.jl's bound for ComplexF64 (rtol 5 eps): |A[:,rd] - A[:,sk] T| < 100 rtol |A|,
    and an exact rank-13 matrix gives k = 13 at rtol 1e-12."""
    rng = np.random.default_rng(1)
    A = orc.fourier(rng.random(128), rng.random(64))   # test/id.jl:7
    nrm = np.linalg.norm(A)
    for trans in ("n", "c"):
        V = orc.zidfact(trans, A, zrand=_ZRand(3), sketch=sk)
        X = A if trans == "n" else A.conj().T
        E = X[:, V.rd - 1] - X[:, V.sk - 1] @ V.T
        assert np.linalg.norm(E) < 100 * 5 * np.finfo(float).eps * nrm, (sk, trans, V.k)
    B = _zlowrank(90, 70, 13, 4)
    V = orc.zidfact("n", B, zrand=_ZRand(4), sketch=sk, rtol=1e-12)
    assert V.k == 13
