"""The allocation-free entry point lra_snormdiff_id_noalloc against lra_snormdiff_id (the
same power iteration on the same kernels) through the C ABI: two calls in a row, the second
on a smaller input, each equal to the allocating entry point's result."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(lra, name, A, V, seed):
    L = lra._lib.load()
    ctx = lra.context()
    m, n = A.shape
    T = V.T if V.T.numel() else lra.empty_colmajor(max(V.k, 1), 1)
    opts = lra.LRAOptions(seed=seed, snorm_niter=40, rtol=1e-12).to_c()
    res = C.c_double(-1.0)
    st = getattr(L, name)(ctx.h, m, n, A.data_ptr(), A.stride(1), V.k, V.sk.data_ptr(),
                          V.rd.data_ptr(), T.data_ptr(), max(1, T.stride(1)), C.byref(opts),
                          None, C.byref(res))
    lra.check(st, name)
    return res.value


@pytest.mark.parametrize("first,second", [((1500, 1200), (400, 300))])
def test_noalloc_matches_snormdiff_id(gpu, first, second):
    import torch
    rng = np.random.default_rng(3)
    for i, (m, n) in enumerate((first, second)):
        print(f"noalloc snormdiff_id call {i}: {m}x{n}", flush=True)
        sig = np.exp(-np.arange(min(m, n)) / 8.0)
        U, _ = np.linalg.qr(rng.standard_normal((m, len(sig))))
        W, _ = np.linalg.qr(rng.standard_normal((n, len(sig))))
        A = gpu.to_device(np.asfortranarray((U * sig) @ W.T))
        V = gpu.idfact(A, seed=i, rtol=1e-8)
        assert 0 < V.k < n  # a rank-deficient ID: its error is not zero
        e_ref = _call(gpu, "lra_snormdiff_id", A, V, 7 + i)
        e_na = _call(gpu, "lra_snormdiff_id_noalloc", A, V, 7 + i)
        torch.cuda.synchronize()
        assert e_ref > 0
        assert e_na == e_ref, (e_na, e_ref)


def test_noalloc_rejects_oversize(gpu):
    import torch
    L = gpu._lib.load()
    ctx = gpu.context()
    A = torch.zeros(8, 8, dtype=torch.float64, device="cuda")
    sk = torch.ones(1, dtype=torch.int64, device="cuda")
    opts = gpu.LRAOptions().to_c()
    res = C.c_double(0.0)
    # 5 Mi rows exceed the 4 Mi-double scratch before any memory is touched
    st = L.lra_snormdiff_id_noalloc(ctx.h, 5 << 20, 1, A.data_ptr(), 5 << 20, 0, sk.data_ptr(),
                                    sk.data_ptr(), A.data_ptr(), 1, C.byref(opts), None,
                                    C.byref(res))
    assert st == gpu._lib.E_UNSUPPORTED
    assert b"maximum" in L.lra_last_error()
