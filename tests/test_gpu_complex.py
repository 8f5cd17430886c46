"""ComplexF64 / ComplexF32 idfact (src/id.jl:434-447 for T <: Complex) on the GPU against the
oracle (the reference's geqp3_adap loop over OpenBLAS zlaqps, the complex sketches' closed
forms) with the GPU's random draws injected, plus test/id.jl's bounds and product identities
for the complex types."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps
EPS32 = float(np.finfo(np.float32).eps)
KINDS = ["randn", "srft", "sub", "sprn"]


def host(t):
    return t.detach().cpu().numpy()


class DeviceZRand:
    """The GPU's complex sketch draws of round rnd (Philox stream rnd under seed), as dense
    sketch matrices for the oracle."""

    def __init__(self, lra, orc, seed):
        self.lra, self.orc, self.seed = lra, orc, seed

    def draws(self, sketch, rnd, m, order):
        d = self.lra.zsketch_draw(sketch, m, order, self.seed, rnd)
        return host(d) if sketch in ("randn", "sub") else tuple(host(t) for t in d)

    def phi(self, sketch, rnd, m, order):
        return self.orc.zsketch_phi(sketch, m, order, self.draws(sketch, rnd, m, order))


def zdecay(m, n, r, rate, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r)))
    return np.asfortranarray((U * rate ** np.arange(r)) @ V.conj().T)


def zpivots_match(g, o, B, tol=1e-10):
    """Pivots equal up to the first step where the two candidates' residual norms (after the
    common earlier pivots are projected out of the sketch B) differ by less than tol of the
    largest column norm; returns the number of equal steps."""
    mx = np.linalg.norm(B, axis=0).max()
    for i, (a, b) in enumerate(zip(g, o)):
        if a != b:
            Q = np.linalg.qr(B[:, np.asarray(o[:i]) - 1])[0] if i else np.zeros((B.shape[0], 0))
            ra = np.linalg.norm(B[:, a - 1] - Q @ (Q.conj().T @ B[:, a - 1]))
            rb = np.linalg.norm(B[:, b - 1] - Q @ (Q.conj().T @ B[:, b - 1]))
            assert abs(ra - rb) < tol * mx, f"pivot {i}: {a} != {b}, gap {abs(ra - rb) / mx:.2e}"
            return i
    return len(g)


def zid_err(A, sk, rd, T, trans="n"):
    X = A if trans == "n" else A.conj().T
    return np.linalg.norm(X[:, rd - 1] - X[:, sk - 1] @ T, 2)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("trans", ["n", "c"])
def test_zsketch_vs_closed_form(gpu, orc, kind, trans):
    import torch
    rng = np.random.default_rng(3)
    m, n, ell = 300, 70, 41
    A = np.asfortranarray(rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
    Ad = gpu.to_device_z(A)
    X = A if trans == "n" else A.conj().T
    P, N = X.shape
    dr = DeviceZRand(gpu, orc, 9)
    B = torch.empty((N, ell), dtype=torch.complex128, device="cuda").t()
    st = gpu._lib.load().lra_zsketch(gpu.context().h, gpu._lib.SKETCH[kind], gpu._lib.TRANS[trans],
                                     m, n, Ad.data_ptr(), m, ell, None, None, 9, 2, B.data_ptr(),
                                     ell)
    gpu.check(st, "lra_zsketch")
    Bo = dr.phi(kind, 2, P, ell) @ X
    assert np.abs(host(B) - Bo).max() <= 1e-12 * np.sqrt(P) * np.abs(A).max()


@pytest.mark.parametrize("fused", ["1", "0"])
def test_zgeqp3_vs_zlaqps_loop(gpu, orc, monkeypatch, fused):
    """The complex rank-adaptive QRCP against the reference's loop over OpenBLAS zlaqps, in
    both forms of the HBM steps (one fused launch per step; two with LRA_QG_FUSED=0)."""
    import torch
    monkeypatch.setenv("LRA_QG_FUSED", fused)
    B = zdecay(120, 400, 60, 0.8, 1)
    Bd = gpu.to_device_z(B).clone()
    n = B.shape[1]
    jp = torch.empty(n, dtype=torch.int64, device="cuda")
    tau = torch.empty(120, dtype=torch.complex128, device="cuda")
    k = C.c_int64(0)
    gpu.check(gpu._lib.load().lra_zgeqp3_adap(gpu.context().h, 120, n, Bd.data_ptr(), 120, -1,
                                              0.0, 1e-10, 32, jp.data_ptr(), tau.data_ptr(),
                                              C.byref(k)), "lra_zgeqp3_adap")
    Bf, p, to, ko = orc.zgeqp3_adap(B, -1, 0.0, 1e-10, 32)
    assert int(k.value) == ko
    g = host(jp)
    i = zpivots_match(g[:ko], p[:ko], B)
    assert i >= 20
    R, Ro = np.triu(host(Bd)[:ko]), np.triu(Bf[:ko])
    assert np.allclose(np.abs(np.diag(R))[:i], np.abs(np.diag(Ro))[:i], rtol=1e-10)


def zgeqp3(gpu, B, rank=-1, rtol=1e-12):
    import torch
    m, n = B.shape
    Bd = gpu.to_device_z(B).clone()
    jp = torch.empty(n, dtype=torch.int64, device="cuda")
    tau = torch.empty(min(m, n), dtype=torch.complex128, device="cuda")
    k = C.c_int64(0)
    gpu.check(gpu._lib.load().lra_zgeqp3_adap(gpu.context().h, m, n, Bd.data_ptr(), m, rank,
                                              0.0, rtol, 32, jp.data_ptr(), tau.data_ptr(),
                                              C.byref(k)), "lra_zgeqp3_adap")
    torch.cuda.synchronize()
    return host(jp), host(tau), int(k.value), host(Bd)


@pytest.mark.parametrize("m,n,rank", [(120, 400, -1), (700, 1500, -1), (1500, 900, 500),
                                      (3000, 1200, -1)])
def test_zgeqp3_fused_matches_two_launch_form(gpu, monkeypatch, m, n, rank):
    """The fused complex HBM steps against the two-launch form on the same graded matrix (no
    near ties): the same rank and pivots, R to rounding, each reflector and tau to rounding
    amplified by the condition of its step, |R(0,0)| / |R(j,j)|.  The shapes take each form
    of the fused kernel (one, two and four waves per column, 8 or 16 rows per lane); rank 500
    ends inside a block."""
    rng = np.random.default_rng(m + n)
    r = min(m, n)
    U = rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))
    V = rng.standard_normal((r, n)) + 1j * rng.standard_normal((r, n))
    B = np.asfortranarray((U * 0.97 ** np.arange(r)) @ V)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("LRA_QG_FUSED", fused)
        print(f"complex fused vs two-launch m={m} n={n} rank={rank} fused={fused}", flush=True)
        out[fused] = zgeqp3(gpu, B, rank)
    (p1, t1, k1, B1), (p0, t0, k0, B0) = out["1"], out["0"]
    assert k1 == k0
    assert np.array_equal(p1, p0)
    kk = min(k1, m, n)
    assert np.abs(np.triu(B1[:kk]) - np.triu(B0[:kk])).max() <= 1e-11 * np.abs(B0[:kk]).max()
    cond = np.abs(B0[0, 0]) / np.abs(np.diag(B0[:kk, :kk]))
    dv = np.abs(np.tril(B1[:, :kk], -1) - np.tril(B0[:, :kk], -1)).max(axis=0)
    assert np.all(dv <= 1e-12 * cond), int(np.argmax(dv / cond))
    assert np.all(np.abs(t1[:kk] - t0[:kk]) <= 1e-12 * cond)


@pytest.mark.parametrize("k,nr,cdiag", [(40, 300, False), (437, 3659, False), (437, 500, True),
                                        (700, 900, False), (1300, 700, True)])
def test_zid_T_vs_scipy(gpu, k, nr, cdiag):
    """lra_zid_T (T = R11 \\ R12, complex) against scipy's triangular solve: a well-conditioned
    upper-triangular R11 with a real diagonal (as the pivoted QR leaves it) or a complex one
    (the solve then scales each row by a unit factor first); k = 700 and 1300 take the real
    2k-row solve in row blocks."""
    import scipy.linalg as sl
    import torch
    rng = np.random.default_rng(k + nr)
    R11 = np.triu(rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))) / np.sqrt(k)
    d = 2.0 + rng.random(k)
    R11[np.diag_indices(k)] = d * np.exp(1j * rng.uniform(0, 2 * np.pi, k)) if cdiag else d
    R12 = rng.standard_normal((k, nr)) + 1j * rng.standard_normal((k, nr))
    R = np.asfortranarray(np.hstack([R11, R12]))
    Rd = gpu.to_device_z(R)
    T = torch.empty((nr, k), dtype=torch.complex128, device="cuda").t()  # column-major k x nr
    print(f"zid_T k={k} nr={nr} complex diagonal={cdiag}", flush=True)
    gpu.check(gpu._lib.load().lra_zid_T(gpu.context().h, k, k + nr, Rd.data_ptr(), k,
                                        T.data_ptr(), k), "lra_zid_T")
    torch.cuda.synchronize()
    Tref = sl.solve_triangular(R11, R12)
    assert np.abs(host(T) - Tref).max() <= 1e-12 * np.abs(Tref).max()


@pytest.mark.parametrize("kind", KINDS + ["none"])
@pytest.mark.parametrize("trans", ["n", "c"])
def test_zidfact_vs_oracle(gpu, orc, kind, trans):
    A = zdecay(700, 500, 120, 0.85, 5)
    print(f"complex idfact sketch={kind} trans={trans}", flush=True)
    V = gpu.idfact(trans, A, seed=4, rtol=1e-10, sketch=kind)
    dr = DeviceZRand(gpu, orc, 4)
    Vo = orc.zidfact(trans, A, zrand=dr, rtol=1e-10, sketch=kind)
    assert V.k == Vo.k and V.rounds == Vo.rounds, (V.k, Vo.k, V.rounds, Vo.rounds)
    X = A if trans == "n" else A.conj().T
    if kind == "none":
        B = X
    else:
        t = 32 * 2 ** (Vo.rounds - 1)
        ell = {"randn": t + 8, "srft": t + 8, "sub": 4 * t + 8, "sprn": t}[kind]
        B = dr.phi(kind, Vo.rounds - 1, X.shape[0], ell) @ X
    assert zpivots_match(host(V.sk), Vo.sk, B) >= min(10, V.k)
    e_g = zid_err(A, host(V.sk), host(V.rd), host(V.T), trans)
    e_o = zid_err(A, Vo.sk, Vo.rd, Vo.T, trans)
    assert e_g <= 10 * max(e_o, 1e-15 * np.linalg.norm(A, 2)), (e_g, e_o)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("kind", KINDS + ["none"])
def test_reference_id_bounds_complex(gpu, orc, dtype, kind):
    """test/id.jl:16-57 for ComplexF64 and ComplexF32: the ID and its products against the
    dense matrices within 100 rtol, rtol = 5 eps(real(T))."""
    rng = np.random.default_rng(0)
    M = orc.fourier(rng.random(128), rng.random(64))
    A = M.astype(dtype)
    rtol = 5 * (EPS if dtype == np.complex128 else EPS32)
    approx = 100 * rtol
    A64 = A.astype(np.complex128)
    nrm = np.linalg.norm(A64)
    for trans in ("n", "c"):
        V = gpu.idfact(trans, A, seed=1, sketch=kind)
        F = gpu.ID.from_matrix(A, V, trans)
        X = A64 if trans == "n" else A64.conj().T
        assert np.linalg.norm(X - host(F.matrix())) < approx * nrm, (trans, V.k)
    V = gpu.idfact(A, seed=2, sketch=kind)
    F = gpu.ID.from_matrix(A, V)
    Vd, Fd = host(V.matrix()), host(F.matrix())
    k, n = Vd.shape
    for N, G, rho in ((Vd, V, k), (Fd, F, A.shape[0])):
        xq = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        xp = rng.standard_normal(rho) + 1j * rng.standard_normal(rho)
        y = N @ xq
        assert np.linalg.norm(y - host(G.apply(xq))) < approx * np.linalg.norm(y)
        y = N.conj().T @ xp
        assert np.linalg.norm(y - host(G.apply(xp, "c"))) < approx * np.linalg.norm(y)
        Bp = rng.standard_normal((rho, rho)) + 1j * rng.standard_normal((rho, rho))
        Bq = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        for Cm, Gm in ((N @ Bq, G.apply(Bq)), (N.conj().T @ Bp, G.apply(Bp, "c")),
                       (Bp @ N, G.rapply(Bp)), (Bq @ N.conj().T, G.rapply(Bq, "c"))):
            assert np.linalg.norm(Cm - host(Gm)) < approx * np.linalg.norm(Cm)


def test_complex_unsupported_options(gpu):
    A = zdecay(100, 80, 20, 0.5, 2)
    with pytest.raises(gpu.LraError):
        gpu.idfact(A, sketch_randn_niter=1)
    with pytest.raises(gpu.LraError):
        gpu.idfact(A, maxdet_tol=0.0)
