"""Matrix-free idfact/snorm (LinearOperator, src/linop.jl) and curfact (src/cur.jl:532-564)
on the GPU against the oracle with the same injected sketches."""
import numpy as np
import pytest

from test_gpu_idfact import DeviceOmega, pivots_match, id_err

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


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


def test_idfact_linop_rejects_structured_sketch(gpu):
    A = np.eye(50)
    L = gpu.LinearOperator.from_matrix(A)
    with pytest.raises(gpu.LraError):
        L.idfact("n", sketch="srft")


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
