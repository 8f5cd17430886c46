"""CPU checks of the boundary: the library loads, exports every symbol include/lra.h
declares, and the host mirror's options carry the reference defaults.  No compute calls."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "lra.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lra_[a-z0-9_]+)\s*\(", txt)) - {"lra_apply_fn"})


def test_library_exports_every_declared_symbol(lra):
    L = lra._lib.load()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_ctypes_signatures_cover_the_header(lra):
    bound = {name for name, _, _ in lra._lib.SIGNATURES}
    assert set(declared_symbols()) <= bound | {"lra_gemm_choice"}


def test_opts_struct_layout_matches_c(lra):
    o = lra._lib.default_opts()   # filled by lra_opts_default in C
    assert o.nb == 32 and o.rank == -1 and o.maxdet_tol == -1 and o.maxdet_niter == -1
    assert o.rtol == 5 * np.finfo(float).eps and o.atol == 0
    assert o.randn_samp_mul == 1 and o.randn_samp_add == 8
    assert o.sub_samp_mul == 4 and o.sub_samp_add == 8
    assert o.snorm_niter == 32 and o.sketch == 1 and o.sketchfact_adap == 1
    assert abs(o.pheig_orthtol - np.sqrt(np.finfo(float).eps)) < 1e-20


def test_python_options_mirror_reference_defaults(lra):
    # src/LowRankApprox.jl:96-119
    o = lra.LRAOptions()
    assert (o.atol, o.maxdet_niter, o.maxdet_tol, o.nb, o.rank, o.sketch, o.sketch_randn_niter,
            o.sketchfact_adap, o.snorm_niter, o.pqrfact_retval) == \
        (0.0, -1, -1.0, 32, -1, "randn", 0, True, 32, "qr")
    c = o.to_c()
    assert c.pqrfact_retval == 3 and c.precision == 0


def test_options_follow_the_element_type(lra):
    """LRAOptions(T) (src/LowRankApprox.jl:96-119): rtol = 5 eps(T), pheig_orthtol =
    sqrt(eps(T)); float32 inputs take Float32's, complex inputs are refused on the host
    before any device work."""
    import pytest
    import torch
    e32 = float(np.finfo(np.float32).eps)
    o = lra.LRAOptions.for_type("f32")
    assert o.rtol == 5 * e32 and abs(o.pheig_orthtol - np.sqrt(e32)) < 1e-20
    assert lra.LRAOptions.for_type("f64") == lra.LRAOptions()
    assert lra.LRAOptions.for_type("f32", rank=7).rank == 7
    assert lra._eltype(np.zeros((2, 2), np.float32)) == "f32"
    assert lra._eltype(torch.zeros(2, 2, dtype=torch.float32)) == "f32"
    assert lra._eltype(np.zeros((2, 2))) == "f64"
    for z in (np.zeros((2, 2), np.complex128), torch.zeros(2, 2, dtype=torch.complex64)):
        with pytest.raises(TypeError):  # never truncated to the real part on a real path
            lra.to_device(z)
    assert lra._eltype(np.zeros((2, 2), np.complex128)) == "c128"
    assert lra._eltype(torch.zeros(2, 2, dtype=torch.complex64)) == "c64"
    assert lra.LRAOptions.for_type("c64").rtol == 5 * e32
    assert lra.LRAOptions.for_type("c128") == lra.LRAOptions()


def test_options_validation(lra):
    import pytest
    for bad in (dict(atol=-1), dict(nb=0), dict(rtol=-1), dict(sketch="foo"),
                dict(precision="fp8")):
        with pytest.raises(ValueError):
            lra.LRAOptions(**bad).check()


def test_no_gpu_means_loud_failure(lra):
    import pytest
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(lra.LraError):
        lra.idfact(np.eye(4))
