"""Float32 inputs (test/id.jl:16-27, test/psvd.jl): LRAOptions(Float32)'s tolerances
(src/LowRankApprox.jl:96-119) with the factorization computed in FP64; the reference test's
error bounds, and the rank of the deterministic path against the oracle at the same rtol.
Complex inputs are rejected."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS32 = float(np.finfo(np.float32).eps)


def host(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def mats(orc):
    rng = np.random.default_rng(0)
    M = orc.fourier_real(rng.random(128), rng.random(64))
    return M, M.astype(np.float32)


def test_float32_default_options(gpu):
    o = gpu.LRAOptions.for_type("f32")
    assert o.rtol == 5 * EPS32
    assert o.pheig_orthtol == pytest.approx(np.sqrt(EPS32))
    assert gpu.LRAOptions.for_type("f64").rtol == gpu.LRAOptions().rtol


@pytest.mark.parametrize("sketch", ["none", "randn", "sub", "srft", "sprn"])
def test_idfact_float32(gpu, mats, sketch):
    """The Float32 loop of test/id.jl: norm(A - Matrix(ID)) < 100 rtol norm(A), both sides."""
    M, A = mats
    rtol = 5 * EPS32
    opts = gpu.LRAOptions.for_type("f32", maxdet_tol=0.0, sketch_randn_niter=1,
                                   sketch=sketch, seed=3)
    print(f"idfact float32 sketch={sketch}", flush=True)
    A64 = A.astype(np.float64)
    nrm = np.linalg.norm(A64, 2)
    for trans, X in (("n", A64), ("c", A64.T)):
        V = gpu.idfact(trans, A, opts)
        F = gpu.ID.from_matrix(A, V, trans)
        Fm = host(F.apply(np.eye(X.shape[1])))
        assert np.linalg.norm(X - Fm, 2) < 100 * rtol * nrm


def test_idfact_float32_defaults_and_rank(gpu, orc, mats):
    """No options: Float32 input takes rtol = 5 eps(Float32), so its rank is that of the
    oracle at that rtol on the same values, and below the Float64 default's."""
    M, A = mats
    A64 = np.asfortranarray(A.astype(np.float64))
    V32 = gpu.idfact(A, sketch="none")
    Vo = orc.idfact("n", A64, sketch="none", rtol=5 * EPS32)
    assert V32.k == Vo.k
    assert np.array_equal(host(V32.sk), np.asarray(Vo.sk))
    V64 = gpu.idfact(A64, sketch="none")
    assert V64.k > V32.k


def test_psvd_float32(gpu, mats):
    M, A = mats
    rtol = 5 * EPS32
    A64 = A.astype(np.float64)
    F = gpu.psvdfact(A, seed=2)
    assert np.linalg.norm(A64 - host(F.matrix()), 2) < 100 * rtol * np.linalg.norm(A64, 2)


def test_complex_on_real_only_paths_rejected(gpu, mats):
    """Complex matrices run through idfact/id; the real-only factorizations refuse them
    rather than dropping the imaginary part."""
    M, _ = mats
    Z = M.astype(np.complex128) * (1 + 0.5j)
    for f in (gpu.pqrfact, gpu.psvdfact, gpu.curfact, gpu.snorm):
        with pytest.raises(TypeError):
            f(Z)
