"""Matrix-free idfact/snorm (LinearOperator, src/linop.jl), curfact (src/cur.jl:532-564) and
the CUR factors (src/cur.jl:85-93) on the GPU against the oracle with the same injected
sketches."""
import numpy as np
import pytest

from test_gpu_idfact import DeviceOmega, pivots_match, id_err

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def host(t):
    return t.detach().cpu().numpy()


class DeviceOmegaBase(DeviceOmega):
    """Device sketches of Philox streams base + round (the second sketchfact of curfact)."""

    def __init__(self, lra, seed, base):
        super().__init__(lra, seed)
        self.base = base

    def gaussian(self, rnd, rows, cols):
        return np.asfortranarray(
            self.lra.gaussian(rows, cols, self.seed, self.base + rnd).cpu().numpy())


@pytest.mark.parametrize("trans", ["n", "c"])
def test_idfact_linop_vs_oracle(gpu, orc, trans):
    """The operator path forms Omega*op(L) as (op(L)' Omega')' through the callback; pivots
    follow the oracle's up to the noise-floor gap rule."""
    A = orc.hilbert(700)[:, :500].copy(order="F")
    L = gpu.LinearOperator.from_matrix(A)
    print(f"linop idfact trans={trans}", flush=True)
    V = L.idfact(trans, seed=4, rtol=1e-12)
    opts = orc.LRAOptions().copy(rtol=1e-12)
    Vo = orc.idfact(trans, A, opts, rand=DeviceOmega(gpu, 4))
    assert V.k == Vo.k and V.rounds == Vo.rounds
    assert pivots_match(V, Vo) >= min(10, V.k)
    e_g = id_err(A, V.sk.cpu().numpy(), V.rd.cpu().numpy(), V.T.cpu().numpy(), trans)
    e_o = id_err(A, Vo.sk, Vo.rd, Vo.T, trans)
    assert e_g <= 10 * max(e_o, 1e-15 * np.linalg.norm(A, 2))


def test_idfact_linop_structured_sketch_uses_randn(gpu):
    """chkopts!(opts, L) (src/LowRankApprox.jl:142-148): a structured sketch on an operator
    warns and runs the Gaussian sketch."""
    A = np.asfortranarray(np.random.default_rng(3).standard_normal((60, 50)))
    L = gpu.LinearOperator.from_matrix(A)
    with pytest.warns(UserWarning, match="randn"):
        V = L.idfact("n", sketch="srft", seed=1, rtol=1e-12)
    W = L.idfact("n", sketch="randn", seed=1, rtol=1e-12)
    assert V.k == W.k
    assert np.array_equal(V.sk.cpu().numpy(), W.sk.cpu().numpy())
    assert np.array_equal(V.T.cpu().numpy(), W.T.cpu().numpy())


def test_linop_failing_operator_reports_status(gpu):
    def bad(X):
        raise RuntimeError("operator failure")
    L = gpu.LinearOperator(40, 30, bad, bad)
    with pytest.raises(gpu.LraError):
        L.idfact("n")


def test_snorm_linop(gpu):
    rng = np.random.default_rng(2)
    A = rng.random((60, 40))
    nrm = np.linalg.norm(A, 2)
    L = gpu.LinearOperator.from_matrix(A)
    assert abs(L.snorm() - nrm) <= 100 * EPS * nrm
    S = A[:40] + A[:40].T
    Ls = gpu.LinearOperator.from_matrix(S)
    assert Ls.hermitian
    assert abs(Ls.snorm() - np.linalg.norm(S, 2)) <= 100 * EPS * np.linalg.norm(S, 2)


def lowrank(m, n, r, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((m, r)) @ rng.standard_normal((r, n)))


@pytest.mark.parametrize("m,n", [(400, 300), (250, 420)])
def test_curfact_vs_oracle(gpu, orc, m, n):
    """Exactly rank-20 matrices: both skeletons and k = 20 match the oracle's."""
    A = lowrank(m, n, 20, m + n)
    U = gpu.curfact(A, seed=6, rtol=1e-10)
    ro, co = orc.curfact(A, rand=DeviceOmega(gpu, 6), rand2=DeviceOmegaBase(gpu, 6, 1 << 32),
                         rtol=1e-10)
    rows, cols = U.rows.cpu().numpy(), U.cols.cpu().numpy()
    assert len(rows) == len(ro) == 20 and len(cols) == len(co) == 20
    assert np.array_equal(rows, ro) and np.array_equal(cols, co)
    # the CUR skeleton reproduces A: A ~ A[:, cols] pinv(A[rows, cols]) A[rows, :]
    C, R, W = A[:, cols - 1], A[rows - 1, :], A[np.ix_(rows - 1, cols - 1)]
    E = A - C @ np.linalg.lstsq(W, R, rcond=None)[0]
    assert np.linalg.norm(E) <= 1e-8 * np.linalg.norm(A)


def test_curfact_symmetric(gpu, orc):
    B = lowrank(300, 300, 15, 9)
    A = np.asfortranarray(B + B.T)
    U = gpu.curfact(A, seed=3, rtol=1e-10)
    ro, co = orc.curfact(A, rand=DeviceOmega(gpu, 3), rtol=1e-10)
    assert np.array_equal(U.rows.cpu().numpy(), U.cols.cpu().numpy())
    assert np.array_equal(U.cols.cpu().numpy(), co) and np.array_equal(ro, co)


def test_sketchfact_matches_idfact_pivots(gpu):
    A = lowrank(300, 200, 12, 1)
    p, k = gpu.sketchfact("n", A, seed=2, rtol=1e-10)
    V = gpu.idfact(A, seed=2, rtol=1e-10)
    assert k == V.k == 12
    assert np.array_equal(p[:k].cpu().numpy(), V.sk.cpu().numpy())


@pytest.mark.parametrize("m,n", [(400, 300), (250, 600)])
def test_cur_build_vs_oracle(gpu, orc, m, n):
    """CUR(A, curfact(A)) (src/cur.jl:85-93): the gathered C and R are A's entries exactly,
    U = pinv(A[rows, cols]) matches the oracle's product V diag(1/sigma) U', and C U R
    approximates A to the ID's accuracy (exactly for the exact-rank part)."""
    rng = np.random.default_rng(m + n)
    k0 = 40
    A = rng.standard_normal((m, k0)) @ np.diag(np.exp(-0.3 * np.arange(k0))) @ \
        rng.standard_normal((k0, n))
    A = np.asfortranarray(A)
    print(f"cur build {m}x{n}", flush=True)
    U = gpu.curfact(A, rtol=1e-9, seed=2)
    F = gpu.CUR.from_packed(A, U)
    rows, cols = host(F.rows), host(F.cols)
    assert np.array_equal(host(F.C), A[:, cols - 1])
    assert np.array_equal(host(F.R), A[rows - 1, :])
    Co, (Vo, so, Uto), Ro = orc.cur_build(A, rows, cols)
    Wp = host(F.U.U) @ np.diag(host(F.U.S)) @ host(F.U.Vt)
    Wpo = Vo @ np.diag(so) @ Uto
    assert np.abs(Wp - Wpo).max() <= 1e-8 * np.abs(Wpo).max()
    X = rng.standard_normal((n, 3))
    Y = host(F.apply(X))
    assert np.abs(Y - Co @ Wpo @ (Ro @ X)).max() <= 1e-8 * np.abs(Y).max()
    Yc = host(F.apply(rng.standard_normal((m, 2)), "c"))
    assert Yc.shape == (n, 2)
    err = np.linalg.norm(A - host(F.C) @ Wp @ host(F.R), 2)
    assert err <= 1e-6 * np.linalg.norm(A, 2)


def test_cur_build_from_skeleton_lists(gpu):
    A = np.asfortranarray(np.random.default_rng(1).standard_normal((50, 40)))
    F = gpu.CUR.from_skeletons(A, [3, 7, 11], [1, 2, 30])
    assert host(F.C).shape == (50, 3) and host(F.R).shape == (3, 40)
    W = A[np.ix_([2, 6, 10], [0, 1, 29])]
    Wp = host(F.U.U) @ np.diag(host(F.U.S)) @ host(F.U.Vt)
    assert np.allclose(Wp, np.linalg.inv(W), rtol=1e-10, atol=1e-12)
