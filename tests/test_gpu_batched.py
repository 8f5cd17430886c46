"""Batched idfact (one persistent fused launch over many blocks): every block against the oracle
with the same injected Omega, against idfact on the block alone, the adaptive fallback, edge
shapes and hipGraph replay."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceOmega:
    """Oracle randomness provider that returns the GPU's Philox sketches."""

    def __init__(self, lra, seed):
        self.lra, self.seed = lra, seed

    def gaussian(self, rnd, rows, cols):
        return np.asfortranarray(self.lra.gaussian(rows, cols, self.seed, rnd).cpu().numpy())


def kernel_block(m, n, rng):
    """K[i, j] = 1 / |x_i - y_j| with x_i uniform in [0,1]^3 and y_j in [3,4] x [0,1]^2."""
    x = rng.random((m, 3))
    y = rng.random((n, 3))
    y[:, 0] += 3.0
    d = np.sqrt(((x[:, None, :] - y[None, :, :]) ** 2).sum(-1))
    return np.asfortranarray(1.0 / d)


def sizes(count, rng, lo=64, hi=1024):
    u = rng.random((count, 2))
    return np.rint(lo * (hi / lo) ** u).astype(int)


def id_err(A, sk, rd, T):
    return np.linalg.norm(A[:, rd - 1] - A[:, sk - 1] @ T, 2) if len(rd) else 0.0


def pivots_match(sk, osk, gap, tol=1e-10):
    for i, (a, b) in enumerate(zip(sk, osk)):
        if a != b:
            assert gap[i] < tol, f"pivot {i}: {a} != {b} with gap {gap[i]:.3e}"
            return i
    return len(sk)


def host(t):
    return t.cpu().numpy()


def test_batched_against_oracle(gpu, orc):
    rng = np.random.default_rng(5)
    shp = [tuple(s) for s in sizes(20, rng)] + [(1500, 300), (700, 1024), (64, 33), (9, 150),
                                                 (333, 1), (1, 70), (257, 45)]
    blocks = [kernel_block(m, n, rng) for m, n in shp]
    print(f"batched vs oracle: {len(blocks)} blocks", flush=True)
    ids = gpu.idfact_batched(blocks, rtol=1e-8, seed=3)
    for A, V, (m, n) in zip(blocks, ids, shp):
        Vo = orc.idfact("n", A, orc.LRAOptions().copy(rtol=1e-8), rand=DeviceOmega(gpu, 3))
        assert V.k == Vo.k, (m, n)
        sk, rd, T = host(V.sk), host(V.rd), host(V.T)
        assert sorted(np.concatenate([sk, rd])) == list(range(1, n + 1))
        pivots_match(sk, Vo.sk, Vo.gap)
        e, eo = id_err(A, sk, rd, T), id_err(A, Vo.sk, Vo.rd, Vo.T)
        assert e <= 10 * max(eo, 1e-14 * np.linalg.norm(A, 2)), (m, n, e, eo)


def test_batched_same_as_alone(gpu, orc):
    """Each block's result in the batch equals idfact on that block alone (same Omega): equal
    rank and pivots up to norm ties, T to the accuracy the conditioning of R11 allows."""
    rng = np.random.default_rng(11)
    shp = [tuple(s) for s in sizes(16, rng)]
    blocks = [kernel_block(m, n, rng) for m, n in shp]
    ids = gpu.idfact_batched(blocks, rtol=1e-8, seed=7)
    for A, V in zip(blocks, ids):
        Va = gpu.idfact(A, rtol=1e-8, seed=7)
        assert V.k == Va.k
        if np.array_equal(host(V.sk), host(Va.sk)):
            assert np.array_equal(host(V.rd), host(Va.rd))
            Tb, Ta = host(V.T), host(Va.T)
            assert np.abs(Tb - Ta).max() <= 1e-6 * max(1.0, np.abs(Ta).max())
        e, ea = id_err(A, host(V.sk), host(V.rd), host(V.T)), id_err(A, host(Va.sk), host(Va.rd),
                                                                     host(Va.T))
        assert e <= 10 * max(ea, 1e-14 * np.linalg.norm(A, 2))


def test_batched_fallback_to_idfact(gpu):
    """A block whose rank reaches the last batched round's target (64) is finished by idfact
    (further adaptive rounds): bit-identical to idfact alone.  Rank 50 is taken by the second
    batched round (order 72)."""
    rng = np.random.default_rng(2)
    full = np.asfortranarray(rng.standard_normal((120, 90)))      # rank 90 > 64
    low = kernel_block(300, 200, rng)
    plan = gpu.BatchedID([full, low, full[:, :50]], rtol=1e-6, seed=1, ldt=90)
    plan.launch()
    ids = plan.finish()
    plan.close()
    assert plan.n_fallback == 1
    for A, V in zip([full, low, full[:, :50]], ids):
        Va = gpu.idfact(A, rtol=1e-6, seed=1)
        assert V.k == Va.k
        assert np.array_equal(host(V.sk), host(Va.sk))
        if V.k == 90:
            assert np.array_equal(host(V.rd), host(Va.rd))
            assert np.array_equal(host(V.T), host(Va.T))
    assert ids[0].k == 90 and ids[2].k == 50


def test_batched_second_round(gpu, orc):
    """Ranks in [32, 64): the second batched round (order 72, two threads per column) for
    n <= 512; wider blocks go to idfact.  Every block against the oracle's two rounds."""
    rng = np.random.default_rng(21)
    shp = [(300, 280), (512, 512), (700, 129), (90, 400), (600, 700), (1000, 64)]
    blocks = [kernel_block(m, n, rng) for m, n in shp]
    plan = gpu.BatchedID(blocks, rtol=1e-13, seed=5)
    plan.launch()
    ids = plan.finish()
    plan.close()
    for A, V, (m, n) in zip(blocks, ids, shp):
        Vo = orc.idfact("n", A, orc.LRAOptions().copy(rtol=1e-13), rand=DeviceOmega(gpu, 5))
        print(f"block {m}x{n}: k={V.k} oracle k={Vo.k} rounds={Vo.rounds}", flush=True)
        assert V.k == Vo.k
        pivots_match(host(V.sk), Vo.sk, Vo.gap)
        e, eo = id_err(A, host(V.sk), host(V.rd), host(V.T)), id_err(A, Vo.sk, Vo.rd, Vo.T)
        assert e <= 10 * max(eo, 1e-15 * np.linalg.norm(A, 2))
    assert any(32 <= V.k < 64 for V in ids)


def test_batched_workgroup_pairs(gpu, orc):
    """n in (512, 1024] with rank in [32, 64): the second round runs on a pair of workgroups
    that swap pivot candidates every step; nothing is left to idfact."""
    rng = np.random.default_rng(33)
    shp = [(800, 1000), (400, 600), (1024, 1024), (700, 513)]
    blocks = [kernel_block(m, n, rng) for m, n in shp]
    plan = gpu.BatchedID(blocks, rtol=1e-10, seed=8)
    plan.launch()
    ids = plan.finish()
    plan.close()
    ks = [V.k for V in ids]
    print(f"pair blocks: k={ks} fallback={plan.n_fallback}", flush=True)
    assert plan.n_fallback == 0
    assert sum(32 <= k < 64 for k in ks) >= 2
    for A, V in zip(blocks, ids):
        Vo = orc.idfact("n", A, orc.LRAOptions().copy(rtol=1e-10), rand=DeviceOmega(gpu, 8))
        assert V.k == Vo.k
        pivots_match(host(V.sk), Vo.sk, Vo.gap)
        e, eo = id_err(A, host(V.sk), host(V.rd), host(V.T)), id_err(A, Vo.sk, Vo.rd, Vo.T)
        assert e <= 10 * max(eo, 1e-15 * np.linalg.norm(A, 2))
        Va = gpu.idfact(A, rtol=1e-10, seed=8)
        assert V.k == Va.k


def test_batched_rank_beyond_ldt(gpu):
    """A fallback block whose rank exceeds the reserved T rows is re-run with its own T."""
    rng = np.random.default_rng(4)
    full = np.asfortranarray(rng.standard_normal((80, 60)))
    ids = gpu.idfact_batched([full], seed=2)
    assert ids[0].k == 60
    Va = gpu.idfact(full, seed=2)
    assert np.array_equal(host(ids[0].sk), host(Va.sk))


def test_batched_edge_shapes(gpu, orc):
    import torch
    rng = np.random.default_rng(8)
    zero = np.zeros((50, 70), order="F")
    odd_ld = gpu.to_device(kernel_block(401, 301, rng))[:399, :]  # leading dimension 401
    shifted = gpu.to_device(kernel_block(400, 120, rng))[1:, :]   # even ld, 8-byte aligned start
    one_col = kernel_block(100, 1, rng)
    blocks = [zero, odd_ld, shifted, one_col]
    ids = gpu.idfact_batched(blocks, rtol=1e-8, seed=4)
    assert ids[0].k == 0 and np.array_equal(host(ids[0].rd), np.arange(1, 71))
    for A, V in zip(blocks[1:], ids[1:]):
        Ah = host(A) if isinstance(A, torch.Tensor) else A
        Ah = np.asfortranarray(Ah)
        Vo = orc.idfact("n", Ah, orc.LRAOptions().copy(rtol=1e-8), rand=DeviceOmega(gpu, 4))
        assert V.k == Vo.k
        pivots_match(host(V.sk), Vo.sk, Vo.gap)
    assert ids[3].k == 1


def test_batched_fixed_rank(gpu, orc):
    """Non-adaptive fixed rank (sketch order rank + 8)."""
    rng = np.random.default_rng(9)
    blocks = [kernel_block(m, n, rng) for m, n in [(300, 200), (90, 500)]]
    ids = gpu.idfact_batched(blocks, rank=12, sketchfact_adap=False, seed=6)
    for A, V in zip(blocks, ids):
        Vo = orc.idfact("n", A, orc.LRAOptions().copy(rank=12, sketchfact_adap=False),
                        rand=DeviceOmega(gpu, 6))
        assert V.k == 12 == Vo.k
        pivots_match(host(V.sk), Vo.sk, Vo.gap)


def test_batched_graph_replay(gpu):
    """The launch captured in a hipGraph and replayed, with eager work and synchronisations
    between replays, gives the eager result every time."""
    import torch
    rng = np.random.default_rng(12)
    blocks = [kernel_block(m, n, rng) for m, n in sizes(40, rng)]
    ref = gpu.idfact_batched(blocks, rtol=1e-8)
    ref = [(host(V.sk), host(V.T)) for V in ref]
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan = gpu.BatchedID(blocks, rtol=1e-8)
        plan.capture()
        for rep in range(3):
            for t in plan.T:
                t.fill_(np.nan)
            plan.launch()
            out = plan.finish()
            x = torch.randn(300, 300, device="cuda", dtype=torch.float64)
            (x @ x).sum().item()
            torch.cuda.synchronize()
            for (sk, T), V in zip(ref, out):
                assert np.array_equal(host(V.sk), sk)
                assert np.array_equal(host(V.T), T), rep
        plan.close()


@pytest.mark.parametrize("w0", ["0,0", "256,128", "512,0", "0,256"])
def test_batched_width_limits_same_bits(gpu, monkeypatch, w0):
    """The autotuner's choice -- which blocks of each round run in the 512-thread kernel and
    which in the 1024-thread one -- changes time only: every forced split of the widths gives
    the bits of the compiled split (ranks, skeletons and T), blocks of every width class and
    both rounds included."""
    rng = np.random.default_rng(21)
    shp = [tuple(s) for s in sizes(36, rng)] + [(800, 1000), (300, 700), (500, 200), (90, 500)]
    blocks = [kernel_block(m, n, rng) for m, n in shp]
    ref = [(V.k, host(V.sk), host(V.T)) for V in gpu.idfact_batched(blocks, rtol=1e-10, seed=4)]
    assert any(k >= 32 for k, _, _ in ref)  # some blocks take the second round
    monkeypatch.setenv("LRA_BATCHED_W0", w0)
    print(f"batched width limits {w0}: {len(blocks)} blocks", flush=True)
    out = gpu.idfact_batched(blocks, rtol=1e-10, seed=4)
    for (k, sk, T), V in zip(ref, out):
        assert V.k == k
        assert np.array_equal(host(V.sk), sk)
        assert np.array_equal(host(V.T), T)


def test_batched_autotuned_plan_same_bits(gpu, monkeypatch):
    """A plan of 256 blocks or more times candidate width limits on its own blocks before it
    returns (launching the batch several times into the caller's buffers): its results equal
    those of the compiled limits (LRA_AUTOTUNE=0), and each limit it keeps is a candidate --
    the compiled one, half of it, or none."""
    rng = np.random.default_rng(44)
    shp = [tuple(s) for s in sizes(300, rng, lo=48, hi=700)]
    blocks = [kernel_block(m, n, rng) for m, n in shp]
    # the GEMM autotuner reads LRA_AUTOTUNE once per process: let a GEMM read it unset first
    gpu.idfact(blocks[0], rtol=1e-9)
    monkeypatch.setenv("LRA_AUTOTUNE", "0")
    plan = gpu.BatchedID(blocks, rtol=1e-9, seed=2)
    compiled = plan.widths
    plan.launch()
    ref = [(V.k, host(V.sk), host(V.T)) for V in plan.finish()]
    plan.close()
    monkeypatch.delenv("LRA_AUTOTUNE")
    plan = gpu.BatchedID(blocks, rtol=1e-9, seed=2)
    widths = plan.widths
    print(f"autotuned batched plan: {len(blocks)} blocks, widths {compiled} -> {widths}",
          flush=True)
    assert len(widths) == len(compiled)
    for w, c in zip(widths, compiled):
        assert w in (c, c // 2, 0)
    plan.launch()
    out = plan.finish()
    plan.close()
    for (k, sk, T), V in zip(ref, out):
        assert V.k == k
        assert np.array_equal(host(V.sk), sk)
        assert np.array_equal(host(V.T), T)
