"""GPU parity of the determinant-maximising swaps (maxdet_swapcols!, src/pqr.jl:445-501)."""
import numpy as np
import pytest

from test_gpu_idfact import run_pair, id_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k,nr,scale,tol", [(20, 80, 2.0, 0.0), (7, 300, 3.0, 0.1),
                                            (64, 1000, 1.5, 0.0), (1, 5, 4.0, 0.0)])
def test_maxdet_swapcols_vs_oracle(gpu, orc, k, nr, scale, tol):
    """Same swaps in the same order, same iteration count; T bit for bit (the update uses the
    separately rounded multiply-add of dger)."""
    import torch
    rng = np.random.default_rng(k * 1000 + nr)
    T = np.asfortranarray(scale * rng.standard_normal((k, nr)))
    p = np.arange(1, k + nr + 1, dtype=np.int64)
    print(f"maxdet k={k} nr={nr}", flush=True)
    po, To, _, ito = orc.maxdet_swapcols(p, T, tol, -1)
    pd = torch.from_numpy(p.copy()).cuda()
    Td = gpu.to_device(T)
    it = gpu.maxdet_swapcols(pd, Td, k, tol, -1)
    assert it == ito and it > 0
    assert np.array_equal(pd.cpu().numpy(), po)
    assert np.array_equal(Td.cpu().numpy(), To)
    assert np.abs(To).max() <= 1 + tol


def test_maxdet_iteration_cap(gpu, orc):
    import torch
    rng = np.random.default_rng(5)
    T = np.asfortranarray(3 * rng.standard_normal((30, 200)))
    p = np.arange(1, 231, dtype=np.int64)
    po, To, _, ito = orc.maxdet_swapcols(p, T, 0.0, 3)
    pd = torch.from_numpy(p.copy()).cuda()
    Td = gpu.to_device(T)
    assert gpu.maxdet_swapcols(pd, Td, 30, 0.0, 3) == ito == 3
    assert np.array_equal(pd.cpu().numpy(), po)
    assert np.array_equal(Td.cpu().numpy(), To)


def test_idfact_with_maxdet(gpu, orc):
    """idfact with maxdet_tol = 0 (the setting of the reference's tests, test/id.jl): the
    swapped skeleton bounds |T| by 1 and the result matches the oracle's."""
    rng = np.random.default_rng(3)
    A = orc.fourier_real(rng.random(300), rng.random(200))
    V, Vo = run_pair(gpu, orc, A, maxdet_tol=0.0, seed=2)
    T = V.T.cpu().numpy()
    assert V.k == Vo.k
    assert np.abs(T).max() <= 1.0 + 1e-12
    e_g = id_err(A, V.sk.cpu().numpy(), V.rd.cpu().numpy(), T)
    e_o = id_err(A, Vo.sk, Vo.rd, Vo.T)
    assert e_g <= 10 * max(e_o, 1e-15 * np.linalg.norm(A, 2))
