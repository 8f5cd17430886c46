"""snorm / snormdiff on the GPU (test/snorm.jl:12-25 bounds) and the ID error through
snormdiff against numpy's 2-norm."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def test_snorm_reference_cases(gpu):
    rng = np.random.default_rng(0)
    rtol = 100 * EPS
    A = rng.random((20, 10))
    nrm = np.linalg.norm(A, 2)
    assert abs(gpu.snorm(A) - nrm) <= rtol * nrm
    A = rng.random((10, 10)); A = A + A.T
    nrm = np.linalg.norm(A, 2)
    assert abs(gpu.snorm(A) - nrm) <= rtol * nrm
    B = rng.random((10, 10))
    nrm = np.linalg.norm(A - B, 2)
    assert abs(gpu.snormdiff(A, B) - nrm) <= rtol * nrm


def test_snorm_large_gap(gpu):
    rng = np.random.default_rng(1)
    U, _ = np.linalg.qr(rng.standard_normal((3000, 50)))
    V, _ = np.linalg.qr(rng.standard_normal((2000, 50)))
    s = np.concatenate([[10.0], np.linspace(5, 1, 49)])  # gap large enough for 32 steps
    A = np.asfortranarray((U * s) @ V.T)
    assert abs(gpu.snorm(A) - 10.0) <= 1e-10


def test_snormdiff_id_matches_dense(gpu, orc):
    A = orc.hilbert(600)
    V = gpu.idfact(A, seed=2, rtol=1e-9)
    E = A[:, V.rd.cpu().numpy() - 1] - A[:, V.sk.cpu().numpy() - 1] @ V.T.cpu().numpy()
    ref = np.linalg.norm(E, 2)
    got = gpu.snormdiff(A, V, snorm_niter=200, rtol=1e-12)
    assert abs(got - ref) <= 1e-3 * ref + 1e-14
