"""GPU parity of the primitives: sketch GEMM (all precisions), QRCP, laqps block, T solve.
Each compares the HIP path (through the C ABI) with the CPU oracle / numpy on the same
inputs."""
import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("M,N,K", [(40, 1024, 1024), (72, 300, 517), (520, 777, 1500),
                                   (17, 33, 5), (1, 1, 1), (130, 257, 64)])
def test_sketch_gemm_f64_left_n(gpu, M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = np.asfortranarray(rng.standard_normal((K, N)))
    Om = np.asfortranarray(rng.standard_normal((M, K)))
    print(f"sketch f64 left n M={M} N={N} K={K}", flush=True)
    B = host(gpu.sketch("left", "n", A, M, omega=Om))
    ref = Om @ A
    assert np.abs(B - ref).max() <= 1e-12 * np.abs(Om).max() * np.abs(A).max() * K


@pytest.mark.parametrize("side,trans", [("left", "c"), ("right", "n"), ("right", "c")])
def test_sketch_gemm_f64_variants(gpu, side, trans):
    rng = np.random.default_rng(7)
    m, n, ell = 300, 200, 40
    A = np.asfortranarray(rng.standard_normal((m, n)))
    X = A if trans == "n" else A.T
    if side == "left":
        Om = np.asfortranarray(rng.standard_normal((ell, X.shape[0])))
        ref = Om @ X
    else:
        Om = np.asfortranarray(rng.standard_normal((X.shape[1], ell)))
        ref = X @ Om
    print(f"sketch f64 {side} {trans}", flush=True)
    B = host(gpu.sketch(side, trans, A, ell, omega=Om))
    assert np.abs(B - ref).max() < 1e-11


@pytest.mark.parametrize("cfg", ["5:1", "6:1", "7:2", "8:1", "9:1", "10:4", "11:1", "12:2", "13:3",
                                 "14:1", "15:2", "16:1", "17:4", "18:3", "18:6", "19:1", "19:4",
                                 "20:2", "21:1", "21:3", "0:1", "4:2"])
@pytest.mark.parametrize("M,N,K", [(40, 1000, 1031), (136, 513, 800), (7, 300, 2)])
def test_sketch_gemm_forced_configs(gpu, monkeypatch, cfg, M, N, K):
    """Every GEMM configuration (the k-contiguous FP64 kernel 5..21 and the general one 0..4),
    forced through LRA_GEMM_CFG, on a generated-Omega sketch with ragged M, N, K (odd K
    exercises the single-element tail of the 16-byte loads)."""
    rng = np.random.default_rng(M * N + K)
    A = np.asfortranarray(rng.standard_normal((K, N)))
    monkeypatch.setenv("LRA_GEMM_CFG", cfg)
    print(f"forced gemm cfg={cfg} M={M} N={N} K={K}", flush=True)
    B = host(gpu.sketch("left", "n", A, M, seed=9, stream=3))
    Om = host(gpu.gaussian(M, K, 9, 3))
    ref = Om @ A
    assert np.abs(B - ref).max() <= 1e-12 * np.abs(Om).max() * np.abs(A).max() * K


@pytest.mark.parametrize("prec,tol", [("f32", 2e-5), ("bf16", 3e-2), ("bf16x3", 2e-5)])
def test_sketch_gemm_lowprec(gpu, prec, tol):
    rng = np.random.default_rng(11)
    K, N, M = 2048, 700, 72
    A = np.asfortranarray(rng.standard_normal((K, N)))
    Om = np.asfortranarray(rng.standard_normal((M, K)))
    print(f"sketch {prec}", flush=True)
    B = host(gpu.sketch("left", "n", A, M, omega=Om, precision=prec))
    ref = Om @ A
    # relative to the scale of a length-K dot product of unit normals
    err = np.abs(B - ref).max() / np.sqrt(K)
    assert err < tol, err


def test_gaussian_fill_stats(gpu):
    G = host(gpu.gaussian(520, 8192, seed=3, stream=4))
    assert abs(G.mean()) < 5e-3 and abs(G.std() - 1) < 5e-3
    G2 = host(gpu.gaussian(520, 8192, seed=3, stream=4))
    assert np.array_equal(G, G2)
    G3 = host(gpu.gaussian(520, 8192, seed=3, stream=5))
    assert not np.array_equal(G, G3)


def graded(m, n, seed, decay=0.9):
    rng = np.random.default_rng(seed)
    U = rng.standard_normal((m, min(m, n)))
    V = rng.standard_normal((min(m, n), n))
    s = decay ** np.arange(min(m, n))
    return np.asfortranarray((U * s) @ V)


def first_mismatch_ok(gpu_p, orc_p, gap, tol=1e-10):
    """Pivots must agree up to the first step whose oracle norm gap (relative to the largest
    column norm) is below tol."""
    for i, (a, b) in enumerate(zip(gpu_p, orc_p)):
        if a != b:
            return gap[i] < tol, i
    return True, len(gpu_p)


@pytest.mark.parametrize("m,n,rtol", [(40, 300, 1e-12), (40, 1024, 0.0), (72, 3000, 1e-10),
                                      (136, 2000, 1e-10), (264, 2500, 1e-13),
                                      (520, 8192, 1e-10), (12, 5, 0.0), (33, 700, 1e-6)])
def test_geqp3_adap_vs_oracle(gpu, orc, m, n, rtol):
    import torch
    B = graded(m, n, m * 7 + n, decay=0.85 if m < 100 else 0.95)
    print(f"geqp3_adap m={m} n={n} rtol={rtol}", flush=True)
    Bf, p_o, tau_o, k_o, gap = orc.geqp3_adap(B, -1, 0.0, rtol, 32, want_gap=True)
    Bd = gpu.to_device(B)
    p_g, tau_g, k_g = gpu.geqp3_adap(Bd, -1, 0.0, rtol, 32)
    torch.cuda.synchronize()
    p_g, tau_g, Bg = host(p_g), host(tau_g), host(Bd)
    ok, i = first_mismatch_ok(p_g, p_o, gap)
    assert ok, f"pivot mismatch at step {i} with gap {gap[i]}"
    if i == len(p_g):
        assert k_g == k_o
        kk = min(k_o, min(m, n))
        R_g, R_o = np.triu(Bg[:kk, :]), np.triu(Bf[:kk, :])
        scale = np.abs(R_o).max()
        assert np.abs(R_g - R_o).max() <= 1e-10 * scale
        # tau = (beta - alpha)/beta is only as accurate as the diagonal: compare where
        # |R_ii| is well above rounding of the trailing updates
        d = np.abs(np.diag(R_o[:, :kk]))
        sel = d > 1e-6 * d[0]
        assert np.allclose(tau_g[:kk][sel], tau_o[:kk][sel], rtol=1e-8, atol=1e-12)


def test_geqp3_rank_cap_and_rank1(gpu, orc):
    x = np.random.default_rng(0).standard_normal(50)
    y = np.random.default_rng(1).standard_normal(400)
    B = np.asfortranarray(np.outer(x, y))
    Bd = gpu.to_device(B)
    p, tau, k = gpu.geqp3_adap(Bd, -1, 0.0, 1e-12, 32)
    assert k == 1
    assert int(host(p)[0]) == int(np.argmax(np.abs(y))) + 1
    B = graded(60, 500, 3)
    Bd = gpu.to_device(B)
    _, _, k = gpu.geqp3_adap(Bd, 20, 0.0, 0.0, 32)
    assert k == 20


@pytest.mark.parametrize("m,n,offset,nb", [(40, 200, 0, 32), (60, 300, 10, 16),
                                           (300, 900, 5, 32)])
def test_laqps_block_vs_oracle(gpu, orc, m, n, offset, nb):
    import torch
    A = graded(m, n, m + n + offset, decay=0.8)
    norms = np.linalg.norm(A[offset:, :], axis=0)
    jp = np.arange(1, n + 1, dtype=np.int64)
    A_o, jp_o, tau_o, vn1_o, vn2_o, kb_o = orc.laqps(A, offset, nb, jp, norms, norms)
    A_ob, jp_ob, tau_ob, _, _, kb_ob = orc.laqps(A, offset, nb, jp, norms, norms,
                                                 use_openblas=True)
    assert kb_o == kb_ob and np.array_equal(jp_o, jp_ob)  # oracle pinned to OpenBLAS
    Ad = gpu.to_device(A)
    jpd = torch.from_numpy(jp).cuda()
    v1 = torch.from_numpy(norms.copy()).cuda()
    v2 = torch.from_numpy(norms.copy()).cuda()
    print(f"laqps m={m} n={n} offset={offset} nb={nb}", flush=True)
    tau_g, kb_g = gpu.laqps(Ad, offset, nb, jpd, v1, v2)
    torch.cuda.synchronize()
    assert kb_g == kb_o
    assert np.array_equal(host(jpd), jp_o)
    assert np.abs(host(Ad) - A_o).max() <= 1e-11 * np.abs(A).max()
    assert np.allclose(host(tau_g), tau_o[:kb_o], rtol=1e-9, atol=1e-13)
    assert np.allclose(host(v1)[kb_o:], vn1_o[kb_o:], rtol=1e-7)


@pytest.mark.parametrize("k,n", [(26, 1024), (426, 8192), (1, 5), (300, 4096), (64, 65),
                                 (15, 48), (17, 50), (16, 16 + 33), (436, 8192), (1100, 1500)])
def test_id_T_vs_scipy(gpu, k, n):
    rng = np.random.default_rng(k + n)
    R = np.triu(rng.standard_normal((k, n)))
    R[np.arange(k), np.arange(k)] = 3.0 + np.abs(R[np.arange(k), np.arange(k)])
    R = np.asfortranarray(R)
    ref = sla.solve_triangular(R[:, :k], R[:, k:], lower=False)
    T = host(gpu.id_T(gpu.to_device(R), k))
    # both solves are backward stable: compare residuals, and the solutions to cond * eps
    res = np.abs(R[:, :k] @ T - R[:, k:]).max()
    assert res <= 1e-13 * np.abs(R).max() * max(1.0, np.abs(ref).max()) * k
    cond = np.linalg.cond(R[:, :k])
    assert np.abs(T - ref).max() <= 1e-14 * cond * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("k,n", [(1185, 1300), (2253, 3758), (2400, 2433), (3600, 3700)])
def test_id_T_blocked_vs_scipy(gpu, k, n):
    """k beyond one MFMA slab (1184): the blocked solve (a GEMM update of each row block's
    right-hand side from the blocks below, then its MFMA solve in place) against scipy's
    dtrsm, on a well-conditioned R (off-diagonal entries of size 1/sqrt(k))."""
    rng = np.random.default_rng(k + n)
    R = np.triu(rng.standard_normal((k, n))) / np.sqrt(k)
    R[np.arange(k), np.arange(k)] = 1.0 + np.abs(R[np.arange(k), np.arange(k)])
    R = np.asfortranarray(R)
    ref = sla.solve_triangular(R[:, :k], R[:, k:], lower=False)
    print(f"blocked id_T k={k} n={n}", flush=True)
    T = host(gpu.id_T(gpu.to_device(R), k))
    res = np.abs(R[:, :k] @ T - R[:, k:]).max()
    assert res <= 1e-13 * max(1.0, np.abs(ref).max()) * np.sqrt(k)
    assert np.abs(T - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("M,N,K", [(32, 32, 8192), (32, 400, 8160), (17, 1000, 1031),
                                   (1, 5, 256), (30, 33, 4099)])
@pytest.mark.parametrize("beta", [0.0, -0.5])
def test_dgemm_tn_skinny(gpu, M, N, K, beta):
    """C = alpha A'B + beta C with M <= 32 and a long K (the blocked QR's V'V and V'X): the
    split-K MFMA kernel against numpy, ragged M, N, K and odd leading dimensions."""
    import torch
    L = gpu._lib.load()
    ctx = gpu.context()
    rng = np.random.default_rng(M * 7 + N + K)
    lda, ldb, ldc = K + 3, K + 1, M + 5
    A = rng.standard_normal((lda, M))
    B = rng.standard_normal((ldb, N))
    C0 = rng.standard_normal((ldc, N))
    Ad, Bd, Cd = gpu.to_device(A), gpu.to_device(B), gpu.to_device(C0)
    print(f"dgemm tn M={M} N={N} K={K} beta={beta}", flush=True)
    gpu.check(L.lra_dgemm(ctx.h, 1, 0, M, N, K, 1.5, Ad.data_ptr(), lda, Bd.data_ptr(), ldb,
                          beta, Cd.data_ptr(), ldc), "dgemm")
    got = host(Cd)
    ref = C0.copy()
    ref[:M, :] = 1.5 * A[:K, :].T @ B[:K, :] + beta * C0[:M, :]
    assert np.abs(got[:M] - ref[:M]).max() <= 1e-13 * K
    assert np.array_equal(got[M:], C0[M:])  # rows past M untouched


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("K,beta", [(32, -1.0), (7, 0.0), (64, 0.25)])
def test_dgemm_short_k_tall_c(gpu, ta, tb, K, beta):
    """C = alpha op(A) op(B) + beta C with a short K and a tall C (the blocked QR's trailing
    update X -= V W): computed as the transposed product with a transposed store."""
    L = gpu._lib.load()
    ctx = gpu.context()
    rng = np.random.default_rng(K * 4 + ta * 2 + tb)
    M, N = 1000, 300
    Am = rng.standard_normal((K, M) if ta else (M, K))
    Bm = rng.standard_normal((N, K) if tb else (K, N))
    lda, ldb, ldc = Am.shape[0] + 1, Bm.shape[0] + 2, M + 3
    A = np.zeros((lda, Am.shape[1])); A[:Am.shape[0]] = Am
    B = np.zeros((ldb, Bm.shape[1])); B[:Bm.shape[0]] = Bm
    C0 = rng.standard_normal((ldc, N))
    Ad, Bd, Cd = gpu.to_device(A), gpu.to_device(B), gpu.to_device(C0)
    print(f"dgemm short-K ta={ta} tb={tb} K={K} beta={beta}", flush=True)
    gpu.check(L.lra_dgemm(ctx.h, ta, tb, M, N, K, -0.75, Ad.data_ptr(), lda, Bd.data_ptr(), ldb,
                          beta, Cd.data_ptr(), ldc), "dgemm")
    got = host(Cd)
    opA = Am.T if ta else Am
    opB = Bm.T if tb else Bm
    ref = C0.copy()
    ref[:M] = -0.75 * opA @ opB + beta * C0[:M]
    assert np.abs(got[:M] - ref[:M]).max() <= 1e-13 * max(K, 1) * 4
    assert np.array_equal(got[M:], C0[M:])


@pytest.mark.parametrize("cfg", ["5:1", "6:2", "7:1", "8:1", "9:1", "10:4", "11:1", "12:2", "13:3",
                                 "14:1", "15:2", "16:1", "17:4", "18:4", "18:6", "19:2", "20:1",
                                 "21:2"])
@pytest.mark.parametrize("M,N,K", [(40, 1000, 1031), (136, 513, 800)])
def test_sketch_gemm_f32_kc_configs(gpu, monkeypatch, cfg, M, N, K):
    """The k-contiguous kernel in its F32 form (FP64 loads rounded to FP32 into LDS,
    v_mfma_f32_16x16x4f32), every configuration forced, against the FP64 product: the error
    is that of FP32 rounding and accumulation."""
    rng = np.random.default_rng(M * N + K + 1)
    A = np.asfortranarray(rng.standard_normal((K, N)))
    monkeypatch.setenv("LRA_GEMM_CFG", cfg)
    print(f"forced f32 kc gemm cfg={cfg} M={M} N={N} K={K}", flush=True)
    B = host(gpu.sketch("left", "n", A, M, seed=9, stream=3, precision="f32"))
    Om = host(gpu.gaussian(M, K, 9, 3))
    ref = Om @ A
    err = np.abs(B - ref).max() / np.sqrt(K)
    assert err < 2e-5, err
