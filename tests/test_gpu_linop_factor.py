"""Matrix-free factorizations (AbstractLinearOperator arguments of src/id.jl:434-447,
src/pqr.jl:285-320, src/psvd.jl:238-290, src/pheig.jl:276-314, src/cur.jl:527-564) against
the oracle on the dense matrix behind the operator, with the same injected sketches: the
power iterations run through L and L' (src/sketch.jl:129-174), skeleton columns come from
one-hot applications (src/util.jl:30-34, src/linop.jl:54-74).  Also the curfact buffer
bound at rtol = 0."""
import ctypes as C

import numpy as np
import pytest

from test_gpu_idfact import DeviceOmega, pivots_match, id_err
from test_gpu_linop_cur import DeviceOmegaBase, lowrank

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


def decaying(m, n, sig, seed):
    rng = np.random.default_rng(seed)
    r = len(sig)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return np.asfortranarray((U * sig) @ V.T)


class CountingOp:
    """A LinearOperator over a dense matrix that counts its applications."""

    def __init__(self, lra, A, hermitian=False):
        import torch
        self.Ad = torch.as_tensor(np.asfortranarray(A), device="cuda")
        self.calls = {"n": 0, "c": 0}

        def mul(X):
            self.calls["n"] += 1
            return self.Ad @ X

        def mulc(X):
            self.calls["c"] += 1
            return self.Ad.t() @ X
        self.L = lra.LinearOperator(A.shape[0], A.shape[1], mul, mulc, hermitian)


@pytest.mark.parametrize("trans", ["n", "c"])
def test_idfact_linop_power_iteration(gpu, orc, trans):
    """sketch_randn_niter = 1 (the setting of every reference test) through the operator."""
    sig = 10.0 ** (-np.arange(80) / 10.0)
    A = decaying(500, 380, sig, 21)
    op = CountingOp(gpu, A)
    print(f"linop idfact niter=1 trans={trans}", flush=True)
    V = op.L.idfact(trans, seed=5, rtol=1e-7, sketch_randn_niter=1)
    # one sketch application and two power-iteration applications per round
    assert op.calls["n"] + op.calls["c"] == 3 * V.rounds
    Vo = orc.idfact(trans, A, orc.LRAOptions().copy(rtol=1e-7, sketch_randn_niter=1),
                    rand=DeviceOmega(gpu, 5))
    assert V.k == Vo.k and V.rounds == Vo.rounds
    assert pivots_match(V, Vo) >= min(10, V.k)
    e_g = id_err(A, host(V.sk), host(V.rd), host(V.T), trans)
    e_o = id_err(A, Vo.sk, Vo.rd, Vo.T, trans)
    assert e_g <= 10 * max(e_o, 1e-15 * np.linalg.norm(A, 2))
    # without power iterations the sketch differs: the option is not dropped
    op.calls = {"n": 0, "c": 0}
    V0 = op.L.idfact(trans, seed=5, rtol=1e-7)
    assert op.calls["n"] + op.calls["c"] == V0.rounds


def test_idfact_hermitian_linop_power_iteration(gpu, orc):
    """A Hermitian operator skips the second orthonormalisation (src/sketch.jl:146-147)."""
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.standard_normal((300, 60)))
    w = 2.0 ** -np.arange(60) * np.where(np.arange(60) % 3 == 0, -1.0, 1.0)
    A = (Q * w) @ Q.T
    A = np.asfortranarray((A + A.T) / 2)
    op = CountingOp(gpu, A, hermitian=True)
    V = op.L.idfact("n", seed=3, rtol=1e-9, sketch_randn_niter=1)
    assert op.calls["n"] + op.calls["c"] == 2 * V.rounds
    Vo = orc.idfact("n", A, orc.LRAOptions().copy(rtol=1e-9, sketch_randn_niter=1),
                    rand=DeviceOmega(gpu, 3))
    assert V.k == Vo.k
    assert pivots_match(V, Vo) >= min(10, V.k)
    e_g = id_err(A, host(V.sk), host(V.rd), host(V.T))
    e_o = id_err(A, Vo.sk, Vo.rd, Vo.T)
    assert e_g <= 10 * max(e_o, 1e-15 * np.linalg.norm(A, 2))


@pytest.mark.parametrize("trans", ["n", "c"])
def test_pqrfact_linop_vs_oracle(gpu, orc, trans):
    sig = 2.0 ** -np.arange(40)
    A = decaying(600, 400, sig, 3)
    L = gpu.LinearOperator.from_matrix(A)
    print(f"linop pqrfact trans={trans}", flush=True)
    F = gpu.pqrfact(trans, L, seed=5, rtol=1e-9)
    Fo = orc.pqrfact(trans, A, rand=DeviceOmega(gpu, 5), rtol=1e-9)
    Q, R, p = host(F.Q), host(F.R), host(F.p)
    assert Q.shape == Fo.Q.shape and R.shape == Fo.R.shape
    k = Q.shape[1]
    assert np.array_equal(p[:k], Fo.p[:k])
    assert np.abs(Q @ R - Fo.Q @ Fo.R).max() <= 1e-10 * np.abs(A).max()
    X = A if trans == "n" else A.T
    assert np.linalg.norm(X[:, p - 1] - Q @ R, 2) <= 10 * max(
        np.linalg.norm(X[:, Fo.p - 1] - Fo.Q @ Fo.R, 2), 1e-15 * np.linalg.norm(A, 2))
    # sketch = none on an operator takes randn (chkopts!, src/LowRankApprox.jl:142-148)
    with pytest.warns(UserWarning):
        Fn = gpu.pqrfact(trans, L, seed=5, rtol=1e-9, sketch="none")
    assert np.array_equal(host(Fn.p), p)


@pytest.mark.parametrize("m,n", [(700, 500), (400, 900)])
def test_psvd_linop_vs_oracle(gpu, orc, m, n):
    sig = 10.0 ** (-np.arange(60) / 8.0)
    A = decaying(m, n, sig, m + n)
    L = gpu.LinearOperator.from_matrix(A)
    print(f"linop psvd {m}x{n}", flush=True)
    F = gpu.psvdfact(L, seed=2, rtol=1e-6)
    Fo = orc.psvdfact(A, rand=DeviceOmega(gpu, 2), rtol=1e-6)
    U, S, Vt = host(F.U), host(F.S), host(F.Vt)
    assert S.shape == Fo.S.shape
    assert np.allclose(S, Fo.S, rtol=1e-9, atol=1e-14)
    assert np.abs(S - sig[: len(S)]).max() <= 1e-5 * sig[0]
    assert np.linalg.norm((U * S) @ Vt - (Fo.U * Fo.S) @ Fo.Vt, 2) <= 1e-8
    s = host(gpu.psvdvals(L, seed=2, rtol=1e-6))
    assert np.allclose(s, S, rtol=1e-12)


def test_pheig_linop_vs_oracle(gpu, orc):
    rng = np.random.default_rng(7)
    n, r = 500, 40
    Qr, _ = np.linalg.qr(rng.standard_normal((n, r)))
    w = np.concatenate([-(2.0 ** -np.arange(15)), 3.0 * 2.0 ** -np.arange(25)])
    A = np.asfortranarray((Qr * w) @ Qr.T)
    A = np.asfortranarray((A + A.T) / 2)
    L = gpu.LinearOperator.from_matrix(A)
    assert L.hermitian
    print("linop pheig", flush=True)
    F = gpu.pheigfact(L, seed=3, rtol=1e-7)
    Fo = orc.pheigfact(A, rand=DeviceOmega(gpu, 3), rtol=1e-7)
    vals, vecs = host(F.values), host(F.vectors)
    assert vals.shape == Fo.values.shape
    assert np.allclose(vals, Fo.values, rtol=1e-9, atol=1e-12)
    Ar = (vecs * vals) @ vecs.T
    assert np.linalg.norm(Ar - (Fo.vectors * Fo.values) @ Fo.vectors.T, 2) <= 1e-8
    v = host(gpu.pheigvals(L, seed=3, rtol=1e-7))
    assert np.allclose(v, vals, rtol=1e-12)
    with pytest.raises(ValueError):
        gpu.pheigfact(gpu.LinearOperator.from_matrix(A, hermitian=False))


@pytest.mark.parametrize("m,n", [(400, 300), (250, 420)])
def test_curfact_linop_vs_oracle(gpu, orc, m, n):
    A = lowrank(m, n, 20, m + n)
    L = gpu.LinearOperator.from_matrix(A)
    U = gpu.curfact(L, seed=6, rtol=1e-10)
    ro, co = orc.curfact(A, rand=DeviceOmega(gpu, 6), rand2=DeviceOmegaBase(gpu, 6, 1 << 32),
                         rtol=1e-10)
    assert np.array_equal(host(U.rows), ro) and np.array_equal(host(U.cols), co)


def test_curfact_symmetric_linop(gpu, orc):
    B = lowrank(300, 300, 15, 9)
    A = np.asfortranarray(B + B.T)
    U = gpu.curfact(gpu.LinearOperator.from_matrix(A), seed=3, rtol=1e-10)
    ro, co = orc.curfact(A, rand=DeviceOmega(gpu, 3), rtol=1e-10)
    assert np.array_equal(host(U.rows), host(U.cols)) and np.array_equal(host(U.cols), co)


def test_id_from_linop(gpu, orc):
    """ID(trans, L, V) takes the skeleton columns from one-hot applications."""
    A = decaying(200, 150, 2.0 ** -np.arange(30), 8)
    L = gpu.LinearOperator.from_matrix(A)
    V = gpu.idfact(L, seed=1, rtol=1e-8)
    F = gpu.ID.from_matrix(L, V)
    assert np.array_equal(host(F.C), A[:, host(V.sk) - 1])


@pytest.mark.parametrize("m,n", [(100, 50), (50, 100)])
def test_curfact_rtol0_stays_in_bounds(gpu, m, n):
    """A full-rank matrix at rtol = 0: the first sketchfact's skeleton can be longer than
    min(m, n); the caller's buffers of min(m, n) entries receive only min(kr, kc) <= min(m, n)
    indices and the memory after them is untouched."""
    import torch
    A = np.asfortranarray(np.random.default_rng(m).standard_normal((m, n)))
    Ad = gpu.to_device(A)
    mn = min(m, n)
    guard = 64
    rows = torch.full((mn + guard,), -7, dtype=torch.int64, device="cuda")
    cols = torch.full((mn + guard,), -7, dtype=torch.int64, device="cuda")
    opts = gpu.LRAOptions(rtol=0.0, seed=1).to_c()
    k = C.c_int64(0)
    ctx = gpu.context()
    print(f"curfact rtol=0 {m}x{n}", flush=True)
    st = gpu._lib.load().lra_curfact(ctx.h, m, n, Ad.data_ptr(), m, C.byref(opts),
                                     rows.data_ptr(), cols.data_ptr(), C.byref(k))
    gpu.check(st, "lra_curfact")
    torch.cuda.synchronize()
    k = int(k.value)
    assert 0 < k <= mn
    assert bool((rows[mn:] == -7).all()) and bool((cols[mn:] == -7).all())
    r, c = host(rows[:k]), host(cols[:k])
    assert r.min() >= 1 and r.max() <= m and c.min() >= 1 and c.max() <= n
    assert len(set(r.tolist())) == k and len(set(c.tolist())) == k
