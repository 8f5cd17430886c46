"""prange (src/prange.jl:14-77) against the oracle on the same inputs: the randn path with the
GPU's Philox sketches injected (right side, and both sides for trans b), the deterministic
sketch = none path, and prange_sub's invariants; errors for the forms not built."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


class DeviceOmega:
    """Oracle randomness provider returning the GPU's Philox sketches of streams base + r."""

    def __init__(self, lra, seed, base=0):
        self.lra, self.seed, self.base = lra, seed, base

    def gaussian(self, rnd, rows, cols):
        return np.asfortranarray(
            self.lra.gaussian(rows, cols, self.seed, self.base + rnd).cpu().numpy())


def decaying(m, n, rate, seed):
    rng = np.random.default_rng(seed)
    r = min(m, n)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return np.asfortranarray((U * np.exp(-rate * np.arange(r))) @ V.T)


def check_basis(Q, X, rtol):
    """Orthonormal columns spanning op(A) to the requested tolerance."""
    k = Q.shape[1]
    assert np.abs(Q.T @ Q - np.eye(k)).max() <= 1e-12 * max(1, k)
    res = np.linalg.norm(X - Q @ (Q.T @ X), 2)
    assert res <= 100 * rtol * np.linalg.norm(X, 2) + 1e-13, res


def check_same(Q, Qo, X, rtol, lead):
    """Q against the oracle's basis from the same sketches: the same rank, the same
    projection of the data to the tolerance, and the same leading columns to rounding.  A
    column whose diagonal of R is d * |R_11| moves by about eps / d under rounding-level
    changes of the sketch, so only the `lead` columns with d >= 1e-6 are compared entrywise."""
    assert Q.shape == Qo.shape
    d = np.linalg.norm(Q @ (Q.T @ X) - Qo @ (Qo.T @ X), 2)
    assert d <= 10 * rtol * np.linalg.norm(X, 2) + 1e-13, d
    j = min(lead, Q.shape[1])
    assert np.abs(Q[:, :j] - Qo[:, :j]).max() <= 1e-8, np.abs(Q[:, :j] - Qo[:, :j]).max()


@pytest.mark.parametrize("trans", ["n", "c"])
@pytest.mark.parametrize("m,n,rate", [(700, 500, 0.05), (400, 900, 0.08), (3000, 260, 0.02)])
def test_prange_randn_vs_oracle(gpu, orc, trans, m, n, rate):
    A = decaying(m, n, rate, m + n)
    rtol = 1e-10
    print(f"prange randn trans={trans} {m}x{n}", flush=True)
    Q = host(gpu.prange(trans, A, rtol=rtol, seed=4))
    Qo = orc.prange(trans, A, rtol=rtol, rand=DeviceOmega(gpu, 4))
    X = A if trans == "n" else A.T
    check_basis(Q, X, rtol)
    check_same(Q, Qo, X, rtol, int(np.log(1e6) / rate))


@pytest.mark.parametrize("trans", ["n", "c"])
def test_prange_none_vs_oracle(gpu, orc, trans):
    A = decaying(500, 350, 0.06, 3)
    print(f"prange none trans={trans}", flush=True)
    Q = host(gpu.prange(trans, A, rtol=1e-9, sketch="none"))
    Qo = orc.prange(trans, A, rtol=1e-9, sketch="none")
    X = A if trans == "n" else A.T
    check_basis(Q, X, 1e-9)
    check_same(Q, Qo, X, 1e-9, int(np.log(1e6) / 0.06))


@pytest.mark.parametrize("trans", ["n", "c"])
def test_prange_sub(gpu, trans):
    """prange_sub: QR of the skeleton columns the subset-sketch ID selects."""
    A = decaying(800, 600, 0.1, 5)
    print(f"prange sub trans={trans}", flush=True)
    Q = host(gpu.prange(trans, A, rtol=1e-8, sketch="sub", seed=2))
    X = A if trans == "n" else A.T
    V = gpu.idfact(trans, A, rtol=1e-8, sketch="sub", seed=2)
    assert Q.shape[1] == V.k
    check_basis(Q, X, 1e-8)
    # the basis spans exactly the chosen skeleton columns
    C = X[:, host(V.sk) - 1]
    assert np.linalg.norm(C - Q @ (Q.T @ C), 2) <= 1e-12 * np.linalg.norm(C, 2)


@pytest.mark.parametrize("sketch", ["randn", "none"])
def test_prange_both_vs_oracle(gpu, orc, sketch):
    A = decaying(450, 450, 0.05, 8)
    A = np.asfortranarray(A + 0.3 * decaying(450, 450, 0.2, 9).T)  # not symmetric
    rtol = 1e-10
    print(f"prange b sketch={sketch}", flush=True)
    Q = host(gpu.prange("b", A, rtol=rtol, sketch=sketch, seed=6))
    Qo = orc.prange("b", A, rtol=rtol, sketch=sketch, rand=DeviceOmega(gpu, 6),
                    rand2=DeviceOmega(gpu, 6, 1 << 32))
    assert Q.shape == Qo.shape
    check_basis(Q, A, rtol)
    check_basis(Q, A.T, rtol)
    for X in (A, A.T):
        d = np.linalg.norm(Q @ (Q.T @ X) - Qo @ (Qo.T @ X), 2)
        assert d <= 10 * rtol * np.linalg.norm(X, 2), d


def test_prange_both_symmetric_is_n(gpu):
    A = decaying(300, 300, 0.1, 11)
    A = np.asfortranarray((A + A.T) / 2)
    Qb = host(gpu.prange("b", A, rtol=1e-10, seed=1))
    Qn = host(gpu.prange("n", A, rtol=1e-10, seed=1))
    assert np.array_equal(Qb, Qn)


def test_prange_errors(gpu):
    A = decaying(300, 200, 0.1, 12)
    with pytest.raises(gpu.LraError):
        gpu.prange("b", A)  # not square
    with pytest.raises(gpu.LraError):
        gpu.prange("n", A, sketch="srft")  # no right-side SRFT
    with pytest.raises(ValueError):
        gpu.prange("x", A)


@pytest.mark.parametrize("shape,trans", [((0, 5), "n"), ((7, 0), "n"), ((0, 5), "c"),
                                         ((6, 0), "c")])
def test_prange_empty(gpu, shape, trans):
    A = np.zeros(shape, order="F")
    Q = gpu.prange(trans, A)
    P = shape[0] if trans == "n" else shape[1]
    assert tuple(Q.shape) == (P, 0)
