"""GPU Householder QR (geqrf layout + explicit Q) against LAPACK (scipy) on the same input."""
import numpy as np
import pytest
import scipy.linalg as sla

pytestmark = pytest.mark.gpu


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("m,n", [(300, 40), (1000, 77), (8192, 436), (256, 256), (5, 1),
                                 (70000, 33), (2000, 1)])
def test_qr_vs_lapack(gpu, m, n):
    """R matches dgeqrf's (same reflector convention, so the same diagonal signs) to rounding;
    Q is orthonormal and Q R reproduces A."""
    rng = np.random.default_rng(m + n)
    A = np.asfortranarray(rng.standard_normal((m, n)) * np.logspace(0, -8, n))
    print(f"qr m={m} n={n}", flush=True)
    Q, R = gpu.qr(A)
    Q, R = host(Q), host(R)
    Rl = sla.qr(A, mode="r")[0][:n, :]
    scale = np.abs(Rl).max(axis=0)  # column scales
    assert np.all(np.abs(R - Rl) <= 1e-10 * scale[None, :] + 1e-13 * np.abs(Rl).max())
    assert np.abs(Q.T @ Q - np.eye(n)).max() <= 1e-12 * max(1, np.sqrt(m) / 10)
    assert np.abs(Q @ R - A).max() <= 1e-13 * np.abs(A).max() * np.sqrt(m)


@pytest.mark.parametrize("n,zero", [(20, (7,)), (70, (7, 40))])
def test_qr_rank_deficient_column(gpu, n, zero):
    """Zero columns give tau = 0 (H = I); with n = 70 they sit in panels whose T (dlarft)
    feeds a trailing update and the explicit Q."""
    rng = np.random.default_rng(1)
    A = rng.standard_normal((500, n))
    for j in zero:
        A[:, j] = 0.0
    A = np.asfortranarray(A)
    Q, R = gpu.qr(A)
    Q, R = host(Q), host(R)
    assert np.abs(Q @ R - A).max() <= 1e-12
    assert np.abs(Q.T @ Q - np.eye(n)).max() <= 1e-12
    Rl = sla.qr(A, mode="r")[0][:n, :]
    assert np.abs(R - Rl).max() <= 1e-11 * np.abs(Rl).max()
