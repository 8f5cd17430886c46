"""GPU parity of the structured sketches and of the mixed-precision idfact forms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def host(t):
    return t.detach().cpu().numpy()


def srft_call(lra, A, ell, d, idx, trans="n"):
    import torch
    L = lra._lib.load()
    ctx = lra.context()
    Ad = lra.to_device(A)
    m, n = Ad.shape
    Q = n if trans == "n" else m
    B = lra.empty_colmajor(ell, Q)
    dd = torch.from_numpy(np.ascontiguousarray(d, dtype=np.float64)).cuda()
    ii = torch.from_numpy(np.ascontiguousarray(idx, dtype=np.int64)).cuda()
    lra.check(L.lra_sketch_srft(ctx.h, lra._lib.TRANS[trans], m, n, Ad.data_ptr(),
                                Ad.stride(1), ell, dd.data_ptr(), ii.data_ptr(), 0, 0,
                                B.data_ptr(), max(1, ell)), "srft")
    return host(B)


@pytest.mark.parametrize("m,n,ell", [(4096, 20, 40), (1000, 30, 40), (97, 5, 6), (12, 3, 5),
                                     (2048, 9, 310), (1536, 4, 73), (2048, 5, 17),
                                     (896, 7, 130), (4100, 3, 40), (263936, 3, 600)])
def test_srft_injected_vs_oracle(gpu, orc, m, n, ell):
    rng = np.random.default_rng(m + ell)
    A = np.asfortranarray(rng.standard_normal((m, n)))
    d = np.where(rng.random(m) > 0.5, 1.0, -1.0)
    idx = rng.integers(1, m + 1, size=ell)
    print(f"srft m={m} n={n} ell={ell} l={orc.srft_l(m, ell)}", flush=True)
    B = srft_call(gpu, A, ell, d, idx)
    Bo = orc.srft_left(A, ell, d, idx)
    # the oracle evaluates w^q and the unit-root powers by repeated products, as the reference
    # does, so its own error grows with the mm = m / l products of the accumulation (3e-8
    # at m = 263936, mm = 1031); the closed form bounds the kernel more tightly
    mm = m // orc.srft_l(m, ell)
    assert np.abs(B - Bo).max() <= 1e-11 * np.sqrt(m) * np.abs(A).max() * max(1.0, mm / 256)
    F = np.fft.fft(d[:, None] * A, axis=0)
    ref = np.empty_like(B)
    for i in range(ell):
        f = idx[i - (i % 2)] - 1
        ref[i] = F[f].real if i % 2 == 0 else F[f].imag
    assert np.abs(B - ref).max() <= 1e-13 * np.log2(m) * np.abs(F).max()


def test_srft_transposed(gpu, orc):
    rng = np.random.default_rng(3)
    A = np.asfortranarray(rng.standard_normal((30, 512)))
    d = np.where(rng.random(512) > 0.5, 1.0, -1.0)
    idx = rng.integers(1, 513, size=40)
    B = srft_call(gpu, A, 40, d, idx, trans="c")
    Bo = orc.srft_left(np.asfortranarray(A.T), 40, d, idx)
    assert np.abs(B - Bo).max() <= 1e-11 * np.sqrt(512) * np.abs(A).max()


def test_srft_config3_shape_closed_form(gpu):
    """65536-row columns, ell = 310 -> l = 256, mm = 256 (the tall SRFT workload), checked
    against numpy.fft on a few columns."""
    m, n, ell = 65536, 6, 310
    rng = np.random.default_rng(9)
    A = np.asfortranarray(rng.standard_normal((m, n)))
    d = np.where(rng.random(m) > 0.5, 1.0, -1.0)
    idx = rng.integers(1, m + 1, size=ell)
    B = srft_call(gpu, A, ell, d, idx)
    F = np.fft.fft(d[:, None] * A, axis=0)
    for i in range(0, ell, 2):
        f = idx[i] - 1
        assert np.allclose(B[i], F[f].real, atol=1e-10 * np.sqrt(m))
        if i + 1 < ell:
            assert np.allclose(B[i + 1], F[f].imag, atol=1e-10 * np.sqrt(m))


@pytest.mark.parametrize("m,n,ell", [(700, 90, 37), (16384, 40, 300), (20000, 24, 61)])
def test_subset_and_sprn_injected(gpu, orc, m, n, ell):
    """Subset and sparse-Gaussian sketches with injected draws against the oracle.  The sparse
    sketch stages each column in LDS up to 16384 rows (the first two shapes) and gathers from
    HBM beyond (the third)."""
    import torch
    L = gpu._lib.load()
    ctx = gpu.context()
    rng = np.random.default_rng(5)
    A = np.asfortranarray(rng.standard_normal((m, n)))
    Ad = gpu.to_device(A)
    r = rng.integers(1, m + 1, size=ell)
    B = gpu.empty_colmajor(ell, n)
    rr = torch.from_numpy(r.astype(np.int64)).cuda()
    gpu.check(L.lra_sketch_subset(ctx.h, 0, m, n, Ad.data_ptr(), m, ell, rr.data_ptr(), 0, 0,
                                  B.data_ptr(), ell), "subset")
    assert np.array_equal(host(B), A[r - 1, :])
    perm = (rng.permutation(m) + 1).astype(np.int64)
    s = rng.standard_normal(m)
    pd = torch.from_numpy(perm).cuda()
    sd = torch.from_numpy(s).cuda()
    gpu.check(L.lra_sketch_sprn(ctx.h, 0, m, n, Ad.data_ptr(), m, ell, pd.data_ptr(),
                                sd.data_ptr(), 0, 0, B.data_ptr(), ell), "sprn")
    assert np.allclose(host(B), orc.sprn_left(A, ell, perm, s), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("sk", ["srft", "sub", "sprn"])
def test_idfact_structured_sketch_residual(gpu, orc, sk):
    """The reference's idfact test bound (test/id.jl:20-32) with the GPU sketches."""
    rng = np.random.default_rng(0)
    A = orc.fourier_real(rng.random(128), rng.random(64))
    nrm = np.linalg.norm(A)
    for trans in ("n", "c"):
        V = gpu.idfact(trans, A, sketch=sk, seed=3)
        X = A if trans == "n" else A.T
        E = X[:, host(V.rd) - 1] - X[:, host(V.sk) - 1] @ host(V.T)
        assert np.linalg.norm(E) < 100 * 5 * EPS * nrm, (trans, V.k)


def decaying(m, n, r, sig, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    return np.asfortranarray((U * sig) @ V.T)


@pytest.mark.parametrize("prec", ["f32", "bf16x3", "bf16"])
def test_mixed_below_floor_equals_fp64(gpu, prec):
    """rtol below the precision's noise floor: the decisive round is FP64 with the same Omega,
    so the mixed result is the FP64 drop-in's factorization: the same rounds, rank and
    skeleton.  Its sketch may come from a different GEMM (the drop-in forms two paired rounds'
    sketches in one product), so T agrees to rounding and the ID errors match."""
    r = 300
    A = decaying(2000, 1500, r, 10.0 ** (-12 * np.arange(r) / (r - 1)), 1)
    V = gpu.idfact(A, seed=4, rtol=1e-10)
    Vm = gpu.idfact(A, seed=4, rtol=1e-10, precision=prec)
    assert V.k == Vm.k and V.rounds == Vm.rounds
    assert np.array_equal(host(V.sk), host(Vm.sk))
    e = np.linalg.norm(A[:, host(V.rd) - 1] - A[:, host(V.sk) - 1] @ host(V.T), 2)
    em = np.linalg.norm(A[:, host(Vm.rd) - 1] - A[:, host(Vm.sk) - 1] @ host(Vm.T), 2)
    assert em <= 2 * e + 1e-14 * np.linalg.norm(A, 2), (e, em)


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
def test_mixed_above_floor_refined(gpu, prec):
    """rtol above the floor: low-precision selection, T refined in FP64; the ID error stays
    within 10x of the FP64 drop-in's on the same matrix."""
    r = 200
    sig = np.exp(-np.arange(r) / 12.0)
    A = decaying(1800, 1400, r, sig, 2)
    V = gpu.idfact(A, seed=5, rtol=1e-5)
    Vm = gpu.idfact(A, seed=5, rtol=1e-5, precision=prec)
    err = lambda W: np.linalg.norm(A[:, host(W.rd) - 1] - A[:, host(W.sk) - 1] @ host(W.T), 2)
    e64, emx = err(V), err(Vm)
    assert abs(V.k - Vm.k) <= 3
    assert emx <= 10 * e64, (emx, e64)
