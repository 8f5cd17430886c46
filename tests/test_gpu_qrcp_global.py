"""The global-memory pivoted QR (sketches beyond the register-resident kernel): forced on the
oracle cases, against the register-resident kernel on the same sketch, and end to end on a
wide matrix whose adaptive rounds outgrow the register-resident kernel."""
import numpy as np
import pytest

from test_gpu_kernels import first_mismatch_ok, graded, host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,rtol", [(40, 300, 1e-12), (72, 3000, 1e-10), (264, 2500, 1e-13),
                                      (12, 5, 0.0), (33, 700, 1e-6), (700, 400, 1e-12),
                                      (1032, 3000, 1e-10), (2056, 2500, 1e-12),
                                      (3500, 1200, 1e-12), (6000, 300, 1e-12)])
@pytest.mark.parametrize("fused", ["1", "0"])
def test_global_qrcp_vs_oracle(gpu, orc, monkeypatch, m, n, rtol, fused):
    """Both forms of the HBM path: one fused launch per step (columns in place, permuted at
    the end) and the two launches per step (LRA_QG_FUSED=0)."""
    import torch
    monkeypatch.setenv("LRA_QRCP_GLOBAL", "1")
    monkeypatch.setenv("LRA_QG_FUSED", fused)
    B = graded(m, n, m * 7 + n, decay=0.85 if m < 100 else 0.95)
    print(f"global geqp3_adap m={m} n={n} rtol={rtol} fused={fused}", flush=True)
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
        assert np.abs(R_g - R_o).max() <= 1e-10 * np.abs(R_o).max()


def test_global_matches_register_kernel(gpu, monkeypatch):
    """The same 520 x 8192 sketch through both engines: same rank and pivots."""
    import torch
    B = graded(520, 8192, 11, decay=0.97)
    B1, B2 = gpu.to_device(B), gpu.to_device(B)
    p1, _, k1 = gpu.geqp3_adap(B1, -1, 0.0, 1e-10, 32)
    monkeypatch.setenv("LRA_QRCP_GLOBAL", "1")
    p2, _, k2 = gpu.geqp3_adap(B2, -1, 0.0, 1e-10, 32)
    torch.cuda.synchronize()
    assert k1 == k2
    assert np.array_equal(host(p1)[:k1], host(p2)[:k2])
    R1, R2 = np.triu(host(B1)[:k1]), np.triu(host(B2)[:k2])
    assert np.abs(R1 - R2).max() <= 1e-10 * np.abs(R1).max()


@pytest.mark.parametrize("m,n,rtol,rank", [(2056, 3000, 1e-10, -1), (1032, 2600, 0.0, 700),
                                            (700, 5000, 1e-13, -1), (5000, 900, 1e-12, -1)])
def test_fused_matches_two_launch_form(gpu, monkeypatch, m, n, rtol, rank):
    """The fused HBM steps against the two-launch form on the same sketch: the same rank, the
    same pivots, R to rounding, and each reflector and tau to rounding amplified by the
    condition of its step, |R(0,0)| / |R(j,j)| (the two forms sum norms and dot products in
    different orders; graded columns, so no near ties); rank 700 ends inside a block
    (700 = 21 blocks of 32 + 28 steps)."""
    import torch
    monkeypatch.setenv("LRA_QRCP_GLOBAL", "1")
    B = graded(m, n, m + 3 * n, decay=0.97)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("LRA_QG_FUSED", fused)
        print(f"fused vs two-launch m={m} n={n} rtol={rtol} rank={rank} fused={fused}", flush=True)
        Bd = gpu.to_device(B)
        p, tau, k = gpu.geqp3_adap(Bd, rank, 0.0, rtol, 32)
        torch.cuda.synchronize()
        out[fused] = (host(p), host(tau), k, host(Bd))
    (p1, t1, k1, B1), (p0, t0, k0, B0) = out["1"], out["0"]
    assert k1 == k0
    assert np.array_equal(p1, p0)
    kk = min(k1, m, n)
    scale = np.abs(B0[:kk]).max()
    assert np.abs(np.triu(B1[:kk]) - np.triu(B0[:kk])).max() <= 1e-11 * scale
    cond = np.abs(B0[0, 0]) / np.abs(np.diag(B0[:kk, :kk]))
    dv = np.abs(np.tril(B1[:, :kk], -1) - np.tril(B0[:, :kk], -1)).max(axis=0)
    assert np.all(dv <= 1e-12 * cond), int(np.argmax(dv / cond))
    assert np.all(np.abs(t1[:kk] - t0[:kk]) <= 1e-12 * cond)


def test_wide_matrix_beyond_register_kernel(gpu, orc):
    """400 x 16384 with rank ~380 at rtol 1e-12: the order-264 and order-520 rounds no longer
    fit the register-resident kernel (16384 columns), so they run on the global path; checked
    against the oracle with the same Omegas."""
    rng = np.random.default_rng(3)
    m, n, r = 400, 16384, 390
    U, _ = np.linalg.qr(rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)))
    s = 10.0 ** (-14 * np.arange(r) / (r - 1))
    A = np.asfortranarray((U * s) @ V.T)

    class DeviceOmega:
        def gaussian(self, rnd, rows, cols):
            return np.asfortranarray(gpu.gaussian(rows, cols, 4, rnd).cpu().numpy())

    V_g = gpu.idfact(A, rtol=1e-12, seed=4)
    V_o = orc.idfact("n", A, orc.LRAOptions().copy(rtol=1e-12), rand=DeviceOmega())
    print(f"wide: k={V_g.k} oracle k={V_o.k} rounds={V_g.rounds}", flush=True)
    assert V_g.rounds == V_o.rounds and V_g.rounds >= 4
    assert V_g.k == V_o.k
    first_bad = next((i for i, (a, b) in enumerate(zip(host(V_g.sk), V_o.sk)) if a != b), None)
    if first_bad is not None:
        assert V_o.gap[first_bad] < 1e-10
    e_g = np.linalg.norm(A[:, host(V_g.rd) - 1] - A[:, host(V_g.sk) - 1] @ host(V_g.T), 2)
    e_o = np.linalg.norm(A[:, V_o.rd - 1] - A[:, V_o.sk - 1] @ V_o.T, 2)
    assert e_g <= 10 * max(e_o, 1e-15)


def test_slow_decay_orders_1032_to_2600(gpu, orc):
    """3000 x 2600 with sigma_i = exp(-0.011 (i-1)): at rtol 1e-10 the rank is about 2090, so
    the adaptive rounds run to order 2056 (rejected, k reaches its target 2048) and then the
    whole 2600 columns; the orders 1032, 2056 and 2600 run on the global path.  Checked against
    the oracle with the same Omegas: rounds, rank (within one step: the diagonal falls by 1.1%
    per step near the threshold), pivots up to the first near-tie, error within 10x."""
    rng = np.random.default_rng(8)
    m, n = 3000, 2600
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = np.exp(-0.011 * np.arange(n))
    A = np.asfortranarray((U * s) @ V.T)
    del U, V

    class DeviceOmega:
        def gaussian(self, rnd, rows, cols):
            return np.asfortranarray(gpu.gaussian(rows, cols, 2, rnd).cpu().numpy())

    print("slow decay 3000x2600: gpu idfact", flush=True)
    V_g = gpu.idfact(A, rtol=1e-10, seed=2)
    print(f"gpu k={V_g.k} rounds={V_g.rounds}; oracle", flush=True)
    V_o = orc.idfact("n", A, orc.LRAOptions().copy(rtol=1e-10), rand=DeviceOmega())
    print(f"oracle k={V_o.k} rounds={V_o.rounds}", flush=True)
    assert V_g.rounds == V_o.rounds == 8
    assert abs(V_g.k - V_o.k) <= 1 and 2048 <= V_o.k < n
    sk_g = host(V_g.sk)
    first_bad = next((i for i, (a, b) in enumerate(zip(sk_g, V_o.sk)) if a != b), None)
    if first_bad is not None:
        assert V_o.gap[first_bad] < 1e-10, (first_bad, V_o.gap[first_bad])
    e_g = np.linalg.norm(A[:, host(V_g.rd) - 1] - A[:, sk_g - 1] @ host(V_g.T), 2)
    e_o = np.linalg.norm(A[:, V_o.rd - 1] - A[:, V_o.sk - 1] @ V_o.T, 2)
    assert e_g <= 10 * max(e_o, 1e-15)
    assert e_g <= 1e3 * 1e-10
