"""End-to-end idfact parity: the HIP path and the oracle on the same matrix with the same
injected Gaussian sketches (the oracle receives the device-generated Omega of every round)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceOmega:
    """Oracle randomness provider that returns the GPU's Philox sketches."""

    def __init__(self, lra, seed):
        self.lra, self.seed = lra, seed

    def gaussian(self, rnd, rows, cols):
        return np.asfortranarray(self.lra.gaussian(rows, cols, self.seed, rnd).cpu().numpy())


def run_pair(lra, orc, A, trans="n", **kw):
    seed = kw.pop("seed", 0)
    V = lra.idfact(trans, A, seed=seed, **kw)
    opts = orc.LRAOptions().copy(**kw)
    Vo = orc.idfact(trans, A, opts, rand=DeviceOmega(lra, seed))
    return V, Vo


def pivots_match(V, Vo, tol=1e-10):
    """Skeleton pivots agree at every step before the first one whose oracle norm gap is below
    tol relative to the largest column norm; returns the number of steps compared."""
    g = V.sk.cpu().numpy()
    for i, (a, b) in enumerate(zip(g, Vo.sk)):
        if a != b:
            assert Vo.gap[i] < tol, f"pivot {i}: {a} != {b} with gap {Vo.gap[i]:.3e}"
            return i
    return len(g)


def id_err(A, sk, rd, T, trans="n"):
    X = A if trans == "n" else A.T
    return np.linalg.norm(X[:, rd - 1] - X[:, sk - 1] @ T, 2)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_hilbert_default_rtol(gpu, orc, seed):
    A = orc.hilbert(1024)
    print(f"hilbert idfact seed={seed}", flush=True)
    V, Vo = run_pair(gpu, orc, A, seed=seed)
    sk, rd, T = V.sk.cpu().numpy(), V.rd.cpu().numpy(), V.T.cpu().numpy()
    assert V.k in (26, 27)                       # README.md:107
    assert V.k == Vo.k
    assert pivots_match(V, Vo) >= 20
    e_g, e_o = id_err(A, sk, rd, T), id_err(A, Vo.sk, Vo.rd, Vo.T)
    assert e_g <= 10 * max(e_o, 1e-15 * np.linalg.norm(A, 2))


def test_hilbert_rank20_and_rtol12(gpu, orc):
    A = orc.hilbert(1024)
    V, Vo = run_pair(gpu, orc, A, rank=20)
    assert V.k == 20 and pivots_match(V, Vo) == 20
    V, Vo = run_pair(gpu, orc, A, rtol=1e-12)
    assert V.k == Vo.k and abs(V.k - 22) <= 1     # README.md:122
    pivots_match(V, Vo)


def decaying(m, n, r, sig, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return np.asfortranarray((U * sig) @ V.T)


@pytest.mark.parametrize("trans", ["n", "c"])
def test_decaying_multiround(gpu, orc, trans):
    m, n, r = 1500, 1200, 200
    sig = 10.0 ** (-12 * np.arange(r) / (r - 1))
    A = decaying(m, n, r, sig, 5)
    print(f"decaying idfact trans={trans}", flush=True)
    V, Vo = run_pair(gpu, orc, A, trans=trans, rtol=1e-10, seed=9)
    assert V.rounds == Vo.rounds and V.rounds >= 3
    assert V.k == Vo.k
    pivots_match(V, Vo)
    e_g = id_err(A, V.sk.cpu().numpy(), V.rd.cpu().numpy(), V.T.cpu().numpy(), trans)
    e_o = id_err(A, Vo.sk, Vo.rd, Vo.T, trans)
    assert e_g <= 10 * e_o


def test_sketch_none(gpu, orc):
    A = orc.fourier_real(np.random.default_rng(0).random(128),
                         np.random.default_rng(1).random(64))
    V = gpu.idfact(A, sketch="none")
    Vo = orc.idfact("n", A, sketch="none")
    assert V.k == Vo.k and np.array_equal(V.sk.cpu().numpy(), Vo.sk)
    E = id_err(A, V.sk.cpu().numpy(), V.rd.cpu().numpy(), V.T.cpu().numpy())
    assert E < 100 * 5 * np.finfo(float).eps * np.linalg.norm(A)   # test/id.jl:20-32


@pytest.mark.parametrize("herm", [False, True])
def test_power_iteration_sketch(gpu, orc, herm):
    """sketch_randn_niter = 1 (src/sketch.jl:129-174): orthrows! by the device QR, one pass
    through A' and A (A itself when Hermitian)."""
    rng = np.random.default_rng(11)
    r = 120
    sig = 10.0 ** (-8 * np.arange(r) / (r - 1))
    U, _ = np.linalg.qr(rng.standard_normal((500, r)))
    if herm:
        A = np.asfortranarray((U * sig) @ U.T)
        A = np.asfortranarray((A + A.T) / 2)
    else:
        V, _ = np.linalg.qr(rng.standard_normal((400, r)))
        A = np.asfortranarray((U * sig) @ V.T)
    print(f"power iteration herm={herm}", flush=True)
    V, Vo = run_pair(gpu, orc, A, sketch_randn_niter=1, rtol=1e-6, seed=6)
    assert V.k == Vo.k and V.rounds == Vo.rounds
    assert pivots_match(V, Vo) >= min(20, V.k)
    sk, rd, T = V.sk.cpu().numpy(), V.rd.cpu().numpy(), V.T.cpu().numpy()
    e_g, e_o = id_err(A, sk, rd, T), id_err(A, Vo.sk, Vo.rd, Vo.T)
    assert e_g <= 10 * max(e_o, 1e-15 * np.linalg.norm(A, 2))
