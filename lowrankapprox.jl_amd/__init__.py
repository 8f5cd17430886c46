"""MI355X-native sketch-then-partial-factorize path of LowRankApprox.jl.

Host-side mirror of the reference API (src/LowRankApprox.jl exports) over the C ABI of
liblra_hip.so.  Matrices are torch CUDA tensors in column-major layout (a numpy Fortran array
or any (m, n) tensor with strides (1, ld)); indices are 1-based like the reference.

  LRAOptions            src/LowRankApprox.jl:77-148
  idfact / id           src/id.jl:429-458           -> IDPackedV(sk, rd, T)
  sketch                src/sketch.jl:35-50
  geqp3_adap / laqps    src/pqr.jl:348-418, src/lapack.jl:117-139
  snorm / snormdiff     src/snorm.jl:14-53
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math
import warnings
from typing import Optional

import numpy as np

from . import _lib
from ._lib import LraError, check

__all__ = ["LRAOptions", "IDPackedV", "idfact", "id", "sketch", "sketchfact", "geqp3_adap",
           "laqps", "id_T", "maxdet_swapcols", "gaussian", "to_device", "LraError", "context",
           "LinearOperator", "CURPackedU", "curfact", "cur", "snorm", "snormdiff", "geqrf",
           "orgqr", "qr", "PartialQR", "PQRFactors", "pqrfact", "pqr", "PartialSVD",
           "psvdfact", "psvdvals", "psvd", "PartialHermEigen", "pheigfact", "pheigvals", "pheig", "ID", "dgemm",
           "BatchedID", "idfact_batched", "Server", "prange", "CUR", "sketch_draw",
           "zsketch_draw", "zgemm"]

_EPS = float(np.finfo(np.float64).eps)


def _torch():
    import torch
    return torch


@dataclasses.dataclass
class LRAOptions:
    """LRAOptions with the reference's Float64 defaults (src/LowRankApprox.jl:96-119).
    `sketchfact_*_samp` are affine maps n -> a*n + b given as (a, b).  `precision` (sketch
    GEMM arithmetic: "f64", "f32", "bf16", "bf16x3") and `seed` (Philox key) are MI355X
    extensions."""
    atol: float = 0.0
    maxdet_niter: int = -1
    maxdet_tol: float = -1.0
    nb: int = 32
    pheig_orthtol: float = math.sqrt(_EPS)
    pqrfact_retval: str = "qr"
    rank: int = -1
    rtol: float = 5 * _EPS
    sketch: str = "randn"
    sketch_randn_niter: int = 0
    sketchfact_adap: bool = True
    sketchfact_randn_samp: tuple = (1, 8)
    sketchfact_srft_samp: tuple = (1, 8)
    sketchfact_sub_samp: tuple = (4, 8)
    snorm_niter: int = 32
    verb: bool = True
    precision: str = "f64"
    seed: int = 0

    @classmethod
    def for_type(cls, eltype: str = "f64", **kw) -> "LRAOptions":
        """LRAOptions(T) (src/LowRankApprox.jl:96-119): rtol = 5 eps(T) and
        pheig_orthtol = sqrt(eps(T)) for the element type T ("f64" or "f32")."""
        eps = _EPS if eltype in ("f64", "c128") else float(np.finfo(np.float32).eps)
        return cls(rtol=5 * eps, pheig_orthtol=math.sqrt(eps)).copy(**kw)

    def copy(self, **kw) -> "LRAOptions":
        o = dataclasses.replace(self)
        for k, v in kw.items():
            if not hasattr(o, k):
                raise AttributeError(k)
            setattr(o, k, v)
        return o

    def check(self) -> None:
        """chkopts! (src/LowRankApprox.jl:133-141)."""
        if not self.atol >= 0: raise ValueError("atol")
        if not self.nb > 0: raise ValueError("nb")
        if not self.pheig_orthtol >= 0: raise ValueError("pheig_orthtol")
        if not self.rtol >= 0: raise ValueError("rtol")
        if self.sketch not in _lib.SKETCH: raise ValueError("sketch")
        if self.precision not in _lib.PREC: raise ValueError("precision")
        self.pqrfact_retval = self.pqrfact_retval.lower()

    def to_c(self) -> _lib.lra_opts:
        self.check()
        o = _lib.default_opts()
        o.atol = self.atol
        o.maxdet_niter = self.maxdet_niter
        o.maxdet_tol = self.maxdet_tol
        o.nb = self.nb
        o.pheig_orthtol = self.pheig_orthtol
        rv = 0
        for ch, bit in (("q", _lib.RET_Q), ("r", _lib.RET_R), ("t", _lib.RET_T)):
            if ch in self.pqrfact_retval:
                rv |= bit
        o.pqrfact_retval = rv
        o.rank = self.rank
        o.rtol = self.rtol
        o.sketch = _lib.SKETCH[self.sketch]
        o.sketch_randn_niter = self.sketch_randn_niter
        o.sketchfact_adap = int(bool(self.sketchfact_adap))
        o.randn_samp_mul, o.randn_samp_add = self.sketchfact_randn_samp
        o.srft_samp_mul, o.srft_samp_add = self.sketchfact_srft_samp
        o.sub_samp_mul, o.sub_samp_add = self.sketchfact_sub_samp
        o.snorm_niter = self.snorm_niter
        o.verb = int(bool(self.verb))
        o.precision = _lib.PREC[self.precision]
        o.seed = int(self.seed) & 0xFFFFFFFFFFFFFFFF
        return o


# ----------------------------------------------------------------------------------------
# device plumbing

def context(device: Optional[int] = None) -> _lib.Context:
    torch = _torch()
    if not torch.cuda.is_available():
        raise LraError("no GPU visible: the MI355X path has no CPU fallback")
    dev = torch.cuda.current_device() if device is None else device
    ctx = _lib.Context.get(dev)
    ctx.bind_stream(torch.cuda.current_stream(dev).cuda_stream)
    return ctx


def _eltype(A) -> str:
    """The reference's element type of an input: "f32", "f64", "c64" or "c128".  Float32
    (ComplexF32) inputs take Float32's default tolerances and are factored in FP64
    (ComplexF64) arithmetic."""
    torch = _torch()
    dt = getattr(A, "dtype", None)
    if dt is None:
        return "f64"
    if isinstance(A, np.ndarray):
        if np.iscomplexobj(A):
            return "c64" if A.dtype == np.complex64 else "c128"
        return "f32" if A.dtype == np.float32 else "f64"
    if isinstance(A, torch.Tensor):
        if A.is_complex():
            return "c64" if A.dtype == torch.complex64 else "c128"
        return "f32" if A.dtype == torch.float32 else "f64"
    return "f64"


def _is_complex(A) -> bool:
    return _eltype(A) in ("c64", "c128")


def _real_only(A, what: str) -> None:
    if _is_complex(A):
        raise TypeError(f"{what}: complex element types are supported by idfact/id only")


def to_device(A, dtype=None):
    """Column-major float64 CUDA tensor for A (numpy array or tensor)."""
    torch = _torch()
    if _is_complex(A):  # never silently truncated to its real part
        raise TypeError("complex data on a real path (use idfact/id for complex matrices)")
    if isinstance(A, np.ndarray):
        A = torch.from_numpy(np.asfortranarray(A, dtype=np.float64 if dtype is None else dtype))
    if A.dim() == 1:
        return A.to("cuda", dtype=torch.float64 if dtype is None else dtype).contiguous()
    if A.dtype != torch.float64:
        A = A.to(torch.float64)
    if not A.is_cuda:
        A = A.cuda()
    colmajor = ((A.shape[0] <= 1 or A.stride(0) == 1) and
                (A.shape[1] <= 1 or A.stride(1) >= max(1, A.shape[0])))
    if not colmajor:
        A = A.t().contiguous().t()
    return A


def to_device_z(A):
    """Column-major complex128 CUDA tensor for A (numpy array or tensor)."""
    torch = _torch()
    if isinstance(A, np.ndarray):
        A = torch.from_numpy(np.asfortranarray(A, dtype=np.complex128))
    if A.is_complex():
        A = A.resolve_conj()  # a lazily conjugated view becomes real data
    if A.dim() == 1:
        return A.to("cuda", dtype=torch.complex128).contiguous()
    if A.dtype != torch.complex128:
        A = A.to(torch.complex128)
    if not A.is_cuda:
        A = A.cuda()
    colmajor = ((A.shape[0] <= 1 or A.stride(0) == 1) and
                (A.shape[1] <= 1 or A.stride(1) >= max(1, A.shape[0])))
    if not colmajor:
        A = A.t().contiguous().t()
    return A


def _z_compact(Z):
    """Z as a compact column-major complex matrix (stride(1) == rows)."""
    if Z.dim() == 1:
        Z = Z.reshape(-1, 1)
    if Z.stride(0) == 1 and (Z.shape[1] <= 1 or Z.stride(1) == Z.shape[0]):
        return Z
    return Z.t().contiguous().t()


def _z_real_view(Z):
    """The real 2p x r view of a compact column-major complex p x r matrix (interleaved
    rows)."""
    torch = _torch()
    p, r = Z.shape
    return torch.view_as_real(Z).as_strided((2 * p, r), (1, 2 * max(p, 1)))


def _z_expand(A):
    """The real 2p x 2q matrix of left multiplication by the complex p x q matrix A on
    interleaved vectors: 2 x 2 blocks [Re -Im; Im Re]."""
    torch = _torch()
    p, q = A.shape
    W = torch.empty((2 * q, 2 * p), dtype=torch.float64, device="cuda").t()
    W[0::2, 0::2] = A.real
    W[0::2, 1::2] = -A.imag
    W[1::2, 0::2] = A.imag
    W[1::2, 1::2] = A.real
    return W


def zgemm(A, B):
    """The complex product A B on the library's FP64 MFMA GEMM (one real GEMM of the
    expanded A with B's interleaved real view)."""
    torch = _torch()
    A, B = to_device_z(A), _z_compact(to_device_z(B))
    p, r = A.shape[0], B.shape[1]
    C = torch.empty((r, p), dtype=torch.complex128, device="cuda").t()
    if p and r:
        if A.shape[1] == 0:
            C.zero_()
        else:
            dgemm(_z_expand(A), _z_real_view(B), C=_z_real_view(C))
    return C


def _ld(A) -> int:
    """Leading dimension of a column-major tensor (a single column's stride(1) is arbitrary in
    torch, so it is not used)."""
    if A.dim() == 2 and A.shape[1] > 1:
        return max(1, A.stride(1))
    return max(1, A.shape[0])


def empty_colmajor(m: int, n: int, dtype=None):
    torch = _torch()
    return torch.empty((n, m), dtype=torch.float64 if dtype is None else dtype,
                       device="cuda").t()


def _ptr(t) -> int:
    return 0 if t is None else t.data_ptr()


# ----------------------------------------------------------------------------------------
# primitives

def gaussian(rows: int, cols: int, seed: int = 0, stream: int = 0):
    """The rows x cols standard-normal matrix of (seed, stream) (Philox4x32-10)."""
    ctx = context()
    out = empty_colmajor(rows, cols)
    check(_lib.load().lra_gaussian_fill(ctx.h, seed, stream, rows, cols, _ptr(out),
                                        max(1, rows)), "lra_gaussian_fill")
    return out


def sketch_draw(sketch: str, m: int, order: int, seed: int = 0, stream: int = 0):
    """The random objects of a left sketch of order `order` on an m-row operand, as the
    sketch kernels draw them from (seed, stream): randn -> Omega (order x m); srft -> (signs
    (m), samples (order, 1-based)); sub -> rows (order, 1-based); sprn -> (permutation (m,
    1-based), coefficients (m)) (src/sketch.jl:91-110, 244-330, 338-361, 566-590)."""
    torch = _torch()
    ctx = context()
    L = _lib.load()
    f64 = dict(dtype=torch.float64, device="cuda")
    i64t = dict(dtype=torch.int64, device="cuda")
    if sketch == "randn":
        O = empty_colmajor(order, m)
        check(L.lra_sketch_draw(ctx.h, 1, seed, stream, m, order, _ptr(O), None), "lra_sketch_draw")
        return O
    if sketch == "srft":
        d, idx = torch.empty(max(m, 1), **f64), torch.empty(max(order, 1), **i64t)
        check(L.lra_sketch_draw(ctx.h, 2, seed, stream, m, order, _ptr(d), _ptr(idx)),
              "lra_sketch_draw")
        return d[:m], idx[:order]
    if sketch == "sub":
        r = torch.empty(max(order, 1), **i64t)
        check(L.lra_sketch_draw(ctx.h, 3, seed, stream, m, order, None, _ptr(r)),
              "lra_sketch_draw")
        return r[:order]
    if sketch == "sprn":
        s, r = torch.empty(max(m, 1), **f64), torch.empty(max(m, 1), **i64t)
        check(L.lra_sketch_draw(ctx.h, 4, seed, stream, m, order, _ptr(s), _ptr(r)),
              "lra_sketch_draw")
        return r[:m], s[:m]
    raise ValueError(sketch)


def sketch(side: str, trans: str, A, order: int, opts: LRAOptions = None, omega=None,
           stream: int = 0, **kw):
    """sketch (src/sketch.jl:35-50), Gaussian case: B = Omega*op(A) or op(A)*Omega."""
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    if order < 0:
        raise ValueError("order")
    if side not in _lib.SIDE: raise ValueError("side")
    if trans not in _lib.TRANS: raise ValueError("trans")
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    if opts.sketch != "randn":
        return _sketch_structured(side, trans, A, order, opts, stream, ctx)
    if side == "left":
        B = empty_colmajor(order, n if trans == "n" else m)
    else:
        B = empty_colmajor(m if trans == "n" else n, order)
    om = to_device(omega) if omega is not None else None
    check(_lib.load().lra_sketch_gaussian(
        ctx.h, _lib.SIDE[side], _lib.TRANS[trans], m, n, _ptr(A), _ld(A), order, _ptr(om),
        opts.seed, stream, _lib.PREC[opts.precision], _ptr(B), _ld(B)), "lra_sketch_gaussian")
    return B


def _sketch_structured(side, trans, A, order, opts, stream, ctx):
    torch = _torch()
    if side == "right":
        # A*S = (S'*A')' with the same structure along the other dimension
        Bt = _sketch_structured("left", "c" if trans == "n" else "n", A, order, opts, stream, ctx)
        return Bt.t()
    m, n = A.shape
    B = empty_colmajor(order, n if trans == "n" else m)
    L = _lib.load()
    t = _lib.TRANS[trans]
    if opts.sketch == "srft":
        check(L.lra_sketch_srft(ctx.h, t, m, n, _ptr(A), _ld(A), order, None, None, opts.seed,
                                stream, _ptr(B), _ld(B)), "lra_sketch_srft")
    elif opts.sketch == "sub":
        check(L.lra_sketch_subset(ctx.h, t, m, n, _ptr(A), _ld(A), order, None, opts.seed,
                                  stream, _ptr(B), _ld(B)), "lra_sketch_subset")
    elif opts.sketch == "sprn":
        check(L.lra_sketch_sprn(ctx.h, t, m, n, _ptr(A), _ld(A), order, None, None, opts.seed,
                                stream, _ptr(B), _ld(B)), "lra_sketch_sprn")
    else:
        raise ValueError(opts.sketch)
    return B


def geqp3_adap(B, rank: int = -1, atol: float = 0.0, rtol: float = 5 * _EPS, nb: int = 32):
    """Rank-adaptive pivoted QR in place (geqp3_adap!, src/pqr.jl:348-418).
    Returns (jpvt 1-based, tau, k)."""
    torch = _torch()
    ctx = context()
    if not (B.is_cuda and B.dtype == torch.float64 and B.stride(0) == 1):
        raise ValueError("B must be a column-major float64 CUDA tensor (updated in place)")
    m, n = B.shape
    jpvt = torch.empty(max(n, 1), dtype=torch.int64, device="cuda")
    tau = torch.empty(max(min(m, n), 1), dtype=torch.float64, device="cuda")
    k = _lib.i64(0)
    check(_lib.load().lra_geqp3_adap_f64(ctx.h, m, n, _ptr(B), _ld(B), rank, atol, rtol, nb,
                                         _ptr(jpvt), _ptr(tau), C.byref(k)), "lra_geqp3_adap_f64")
    return jpvt[:n], tau[:min(m, n)], int(k.value)


def laqps(A, offset: int, nb: int, jpvt, vn1, vn2):
    """One ?laqps block step in place (src/lapack.jl:117-139).  Returns (tau, kb)."""
    torch = _torch()
    ctx = context()
    m, n = A.shape
    tau = torch.zeros(max(nb, 1), dtype=torch.float64, device="cuda")
    kb = _lib.i64(0)
    check(_lib.load().lra_laqps_f64(ctx.h, m, n, offset, nb, C.byref(kb), _ptr(A), _ld(A),
                                    _ptr(jpvt), _ptr(tau), _ptr(vn1), _ptr(vn2), None, None,
                                    max(1, n)), "lra_laqps_f64")
    return tau[: int(kb.value)], int(kb.value)


def geqrf(A):
    """Householder QR in place (LAPACK geqrf layout); returns tau.  A: column-major float64
    CUDA tensor, m >= n."""
    torch = _torch()
    ctx = context()
    m, n = A.shape
    tau = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
    check(_lib.load().lra_geqrf_f64(ctx.h, m, n, _ptr(A), _ld(A), _ptr(tau)), "lra_geqrf_f64")
    return tau[:n]


def orgqr(A, tau, k: int = None):
    """Thin Q (m x k) from a geqrf factorization."""
    ctx = context()
    m = A.shape[0]
    k = tau.shape[0] if k is None else k
    Q = empty_colmajor(m, k)
    check(_lib.load().lra_orgqr_f64(ctx.h, m, k, _ptr(A), _ld(A), _ptr(tau), _ptr(Q),
                                    max(1, m)), "lra_orgqr_f64")
    return Q


def qr(A):
    """(Q, R) of a tall matrix (qr! of the reference's factorizations)."""
    torch = _torch()
    X = to_device(A).clone() if not isinstance(A, np.ndarray) else to_device(A)
    X = X.t().contiguous().t() if X.stride(0) != 1 else X
    m, n = X.shape
    tau = geqrf(X)
    Q = orgqr(X, tau, n)
    R = torch.triu(X[:n, :])
    return Q, R


def id_T(R, k: int):
    """T = R[:, :k] \\ R[:, k:] with R upper trapezoidal (maxdet_t, src/pqr.jl:438-442)."""
    ctx = context()
    n = R.shape[1]
    T = empty_colmajor(max(k, 1), max(n - k, 0))
    check(_lib.load().lra_id_T_f64(ctx.h, k, n, _ptr(R), _ld(R), _ptr(T), max(1, k)),
          "lra_id_T_f64")
    return T[:k, :]


def maxdet_swapcols(p, T, k: int, tol: float = 0.0, niter: int = -1) -> int:
    """Determinant-maximising swaps on (p, T) in place (maxdet_swapcols!, src/pqr.jl:445-501).
    p: int64 CUDA vector (n, 1-based), T: column-major k x (n-k) CUDA tensor.  Returns the
    iteration count."""
    ctx = context()
    n = p.shape[0]
    it = _lib.i64(0)
    ldt = _ld(T) if T.numel() else max(1, k)
    check(_lib.load().lra_maxdet_swapcols_f64(ctx.h, k, n, _ptr(p), _ptr(T), ldt, tol, niter,
                                              C.byref(it)), "lra_maxdet_swapcols_f64")
    return int(it.value)


# ----------------------------------------------------------------------------------------
# IDPackedV / idfact

@dataclasses.dataclass
class IDPackedV:
    """IDPackedV (src/id.jl:15-19): A[:, rd] ~ A[:, sk] * T (1-based indices, device)."""
    sk: object
    rd: object
    T: object
    rounds: int = 1

    @property
    def k(self) -> int:
        return int(self.sk.shape[0])

    def __getitem__(self, key):
        # getindex (src/id.jl:72-81)
        torch = _torch()
        if key == "sk": return self.sk
        if key == "rd": return self.rd
        if key == "T": return self.T
        if key == "k": return self.k
        if key == "p": return torch.cat([self.sk, self.rd])
        raise KeyError(key)

    @property
    def n(self) -> int:
        return self.k + int(self.rd.shape[0])

    def apply(self, X, trans: str = "n"):
        """V X (trans "n", X n x nrhs) or V' X (trans "c", X k x nrhs) for V = [I T] P'
        (mul!/mulc! of IDPackedV, src/id.jl:94-152); a vector gives a vector."""
        if self.T.is_complex():
            return self._zapply(X, trans)
        ctx = context()
        X, vec = _in_mat(X)
        nrhs = X.shape[1]
        rows = self.k if trans == "n" else self.n
        Y = empty_colmajor(rows, nrhs)
        ldt = _ld(self.T) if self.T.numel() else max(1, self.k)
        check(_lib.load().lra_id_apply(ctx.h, 0 if trans == "n" else 1, self.k, self.n,
                                       _ptr(self.sk), _ptr(self.rd), _ptr(self.T), ldt, nrhs,
                                       _ptr(X), _ld(X), _ptr(Y), max(1, rows)), "lra_id_apply")
        return _out_mat(Y, vec)

    def _zapply(self, X, trans: str):
        """Complex V X = X[sk] + T X[rd] or V^H X = P [I; T^H] X (src/id.jl:94-152)."""
        torch = _torch()
        X = to_device_z(X)
        vec = X.dim() == 1
        if vec:
            X = X.reshape(-1, 1)
        if trans == "n":
            Y = X[self.sk - 1, :].clone()
            if self.n > self.k:
                Y += zgemm(self.T, X[self.rd - 1, :])
        else:
            Y = torch.empty((self.n, X.shape[1]), dtype=torch.complex128, device="cuda")
            Y[self.sk - 1, :] = X
            if self.n > self.k:
                Y[self.rd - 1, :] = zgemm(self.T.conj().t(), X)
        Y = _z_compact(Y)
        return Y[:, 0] if vec else Y

    def rapply(self, X, trans: str = "n"):
        """X V (trans "n", X m x k) or X V' (trans "c", X m x n): the right products of
        src/id.jl:154-200, as (V' X')' and (V X')'."""
        return _right_apply(self, X, trans)

    def matrix(self):
        """The k x n interpolation matrix [I T] P' (Matrix(V), src/id.jl:59-66): column
        sk[i] is e_i and column rd[j] is T[:, j], scattered on the device."""
        torch = _torch()
        if self.T.is_complex():
            M = torch.zeros((self.n, self.k), dtype=torch.complex128, device="cuda").t()
        else:
            M = empty_colmajor(self.k, self.n)
            M.zero_()
        if self.k:
            M[torch.arange(self.k, device="cuda"), self.sk - 1] = 1.0
            if self.n > self.k:
                M[:, self.rd - 1] = self.T
        return M


def dgemm(A, B, transa: str = "n", transb: str = "n", alpha: float = 1.0, C=None,
          beta: float = 0.0):
    """alpha op(A) op(B) + beta C on the library's FP64 MFMA GEMM."""
    ctx = context()
    A, B = to_device(A), to_device(B)
    ta, tb = int(transa != "n"), int(transb != "n")
    m = A.shape[1] if ta else A.shape[0]
    k = A.shape[0] if ta else A.shape[1]
    n = B.shape[0] if tb else B.shape[1]
    if C is None:
        C = empty_colmajor(m, n)
        beta = 0.0
    check(_lib.load().lra_dgemm(ctx.h, ta, tb, m, n, k, alpha, _ptr(A), _ld(A), _ptr(B), _ld(B),
                                beta, _ptr(C), _ld(C)), "lra_dgemm")
    return C


def _gesvj(X):
    """One-sided Jacobi SVD X = U diag(s) V' of an m x n (m >= n) matrix (X is copied)."""
    torch = _torch()
    ctx = context()
    X0 = to_device(X)
    m, n = X0.shape
    X = empty_colmajor(m, n)
    X.copy_(X0)
    U, V = empty_colmajor(m, n), empty_colmajor(n, n)
    s = torch.empty(n, dtype=torch.float64, device="cuda")
    if n:
        check(_lib.load().lra_gesvj_f64(ctx.h, m, n, _ptr(X), _ld(X), _ptr(U), max(1, m),
                                        _ptr(s), _ptr(V), max(1, n)), "lra_gesvj_f64")
    return U, s, V


def _scale_rows(d, X):
    ctx = context()
    check(_lib.load().lra_scale_rows(ctx.h, X.shape[0], X.shape[1], _ptr(d), _ptr(X), _ld(X)),
          "lra_scale_rows")
    return X


def _in_mat(X):
    """(X on the device as a matrix, whether X was a vector): a vector is one column (the
    reference's methods on AbstractVector, e.g. src/psvd.jl:61-66, are those on a one-column
    matrix, and return a vector)."""
    X = to_device(X)
    return (X.reshape(-1, 1), True) if X.dim() == 1 else (X, False)


def _out_mat(Y, vec: bool):
    return Y[:, 0] if vec else Y


def _as_matrix(X):
    return _in_mat(X)[0]


def _right_apply(F, X, trans: str):
    """X op(F) through F's left product: X op(F) = (op(F)' X')', returned column-major; a
    vector x is the row x' and gives a vector (' is the conjugate transpose)."""
    X = to_device_z(X) if (_is_complex(X) or _factor_is_complex(F)) else to_device(X)
    vec = X.dim() == 1
    if vec:
        X = X.reshape(1, -1)
    Y = F.apply(X.mH, "c" if trans == "n" else "n").mH
    Y = Y.t().contiguous().t()
    return Y[0, :] if vec else Y


def _factor_is_complex(F) -> bool:
    if isinstance(F, IDPackedV):
        return F.T.is_complex()
    if isinstance(F, ID):
        return F.C.is_complex()
    return False


@dataclasses.dataclass
class ID:
    """ID (src/id.jl:212-235): A ~ C V with C = A[:, sk] (m x k) and V = [I T] P'."""
    C: object
    V: IDPackedV

    @classmethod
    def from_matrix(cls, A, V: IDPackedV, trans: str = "n"):
        """ID(trans, A, V) (src/id.jl:219-223): C = getcols(trans, A, sk); A may be a
        LinearOperator (columns from one-hot applications)."""
        if isinstance(A, LinearOperator):
            return cls(to_device(A.getcols(V.sk, trans)), V)
        if _is_complex(A) or V.T.is_complex():
            A = to_device_z(A)
            X = A if trans == "n" else A.conj().t()
            return cls(_z_compact(X[:, V.sk - 1]), V)
        A = to_device(A)
        X = A if trans == "n" else A.t()
        return cls(X[:, V.sk - 1].t().contiguous().t(), V)

    def apply(self, X, trans: str = "n"):
        """A X = C (V X) (trans "n") or A' X = V' (C' X) (trans "c") (src/id.jl:263-300)."""
        if self.C.is_complex():
            X = to_device_z(X)
            vec = X.dim() == 1
            if vec:
                X = X.reshape(-1, 1)
            if trans == "n":
                Y = zgemm(self.C, self.V.apply(X, "n"))
            else:
                Y = self.V.apply(zgemm(self.C.conj().t(), X), "c")
            return Y[:, 0] if vec else Y
        X, vec = _in_mat(X)
        if trans == "n":
            Y = dgemm(self.C, self.V.apply(X, "n"))
        else:
            Y = self.V.apply(dgemm(self.C, X, transa="c"), "c")
        return _out_mat(Y, vec)

    def rapply(self, X, trans: str = "n"):
        """X A = (X C) V (trans "n") or X A' = (X V') C' ("c") (src/id.jl:302-330)."""
        X = to_device_z(X) if self.C.is_complex() else to_device(X)
        vec = X.dim() == 1
        if vec:
            X = X.reshape(1, -1)
        if self.C.is_complex():
            X = to_device_z(X)
            if trans == "n":
                Y = self.V.rapply(zgemm(X, self.C), "n")
            else:
                Y = zgemm(self.V.rapply(X, "c"), self.C.conj().t())
        elif trans == "n":
            Y = self.V.rapply(dgemm(X, self.C), "n")
        else:
            Y = dgemm(self.V.rapply(X, "c"), self.C, transb="c")
        return Y[0, :] if vec else Y

    def matrix(self):
        """Matrix(A) = C V (src/id.jl:244)."""
        if self.C.is_complex():
            return zgemm(self.C, self.V.matrix())
        return dgemm(self.C, self.V.matrix())


def idfact(trans, A=None, opts: LRAOptions = None, **kw) -> IDPackedV:
    """idfact(trans, A, opts; kw...) (src/id.jl:434-447); idfact(A) means trans = :n."""
    torch = _torch()
    if A is None or isinstance(A, LRAOptions):
        trans, A, opts = "n", trans, A
    if isinstance(A, LinearOperator):
        return A.idfact(trans, opts, **kw)
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    opts.pqrfact_retval = "t"
    if _is_complex(A):
        return _zidfact(trans, A, opts)
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    N = n if trans == "n" else m
    sk = torch.empty(max(N, 1), dtype=torch.int64, device="cuda")
    rd = torch.empty(max(N, 1), dtype=torch.int64, device="cuda")
    out = _lib.lra_id_out()
    out.sk, out.rd = _ptr(sk), _ptr(rd)
    out.T, out.ldt = None, 0  # library-owned T, copied out below
    copts = opts.to_c()
    check(_lib.load().lra_idfact(ctx.h, _lib.TRANS[trans], m, n, _ptr(A), _ld(A),
                                 C.byref(copts), C.byref(out)), "lra_idfact")
    k = int(out.k)
    T = _borrow_T(out, k, N - k)
    return IDPackedV(sk[:k], rd[: N - k], T, int(out.rounds))


def _zidfact(trans: str, A, opts: LRAOptions) -> IDPackedV:
    """idfact for ComplexF64 (ComplexF32 inputs are widened) through lra_zidfact."""
    torch = _torch()
    ctx = context()
    A = to_device_z(A)
    m, n = A.shape
    N = n if trans == "n" else m
    sk = torch.empty(max(N, 1), dtype=torch.int64, device="cuda")
    rd = torch.empty(max(N, 1), dtype=torch.int64, device="cuda")
    out = _lib.lra_id_out()
    out.sk, out.rd = _ptr(sk), _ptr(rd)
    out.T, out.ldt = None, 0
    copts = opts.to_c()
    check(_lib.load().lra_zidfact(ctx.h, _lib.TRANS[trans], m, n, _ptr(A), _ld(A),
                                  C.byref(copts), C.byref(out)), "lra_zidfact")
    k = int(out.k)
    nr = N - k
    T = torch.empty((nr, k), dtype=torch.complex128, device="cuda").t()
    if k > 0 and nr > 0:
        # T is complex (k x nr, ldt complex elements): copy it as 2k x nr doubles
        _copy_device(T.data_ptr(), int(out.T), 2 * k, nr, 2 * int(out.ldt), 2 * k)
    return IDPackedV(sk[:k], rd[:nr], T, int(out.rounds))


def zsketch_draw(sketch: str, m: int, order: int, seed: int = 0, stream: int = 0):
    """The random objects of a complex left sketch (lra_zsketch_draw): randn -> Phi
    (order x m complex); srft -> (signs (m complex), samples (order)); sub -> rows (order);
    sprn -> (permutation (m), coefficients (m complex))."""
    torch = _torch()
    ctx = context()
    L = _lib.load()
    c128 = dict(dtype=torch.complex128, device="cuda")
    i64t = dict(dtype=torch.int64, device="cuda")
    if sketch == "randn":
        P = torch.empty((max(m, 1), max(order, 1)), **c128).t()
        check(L.lra_zsketch_draw(ctx.h, 1, seed, stream, m, order, _ptr(P), None),
              "lra_zsketch_draw")
        return P[:order, :m]
    if sketch == "srft":
        d, idx = torch.empty(max(m, 1), **c128), torch.empty(max(order, 1), **i64t)
        check(L.lra_zsketch_draw(ctx.h, 2, seed, stream, m, order, _ptr(d), _ptr(idx)),
              "lra_zsketch_draw")
        return d[:m], idx[:order]
    if sketch == "sub":
        r = torch.empty(max(order, 1), **i64t)
        check(L.lra_zsketch_draw(ctx.h, 3, seed, stream, m, order, None, _ptr(r)),
              "lra_zsketch_draw")
        return r[:order]
    if sketch == "sprn":
        s, r = torch.empty(max(m, 1), **c128), torch.empty(max(m, 1), **i64t)
        check(L.lra_zsketch_draw(ctx.h, 4, seed, stream, m, order, _ptr(s), _ptr(r)),
              "lra_zsketch_draw")
        return r[:m], s[:m]
    raise ValueError(sketch)


class BatchedID:
    """A planned batched idfact of many independent blocks (lra_idfact_batched_plan): every
    block gets the result of idfact on it alone with the same options; the first adaptive
    round of all blocks runs as one persistent fused launch (sketch GEMM + pivoted QR + T),
    and blocks whose rank reaches that round's target are completed by idfact.

    plan = BatchedID(blocks, opts); plan.capture() (optional hipGraph); plan.launch();
    ids = plan.finish()  ->  list of IDPackedV.  `ldt` is the row capacity reserved for each
    T (default: the largest sketch order of the batched rounds, which bounds every rank they
    return).
    """

    def __init__(self, blocks, opts: LRAOptions = None, ldt: Optional[int] = None, **kw):
        torch = _torch()
        first = blocks[0] if len(blocks) else None
        opts = (opts or LRAOptions.for_type(_eltype(first))).copy(**kw)
        opts.pqrfact_retval = "t"
        self.opts = opts
        self.ctx = context()
        self.blocks = [to_device(A) for A in blocks]
        nb = len(self.blocks)
        self.nblk = nb
        self.m = np.array([A.shape[0] for A in self.blocks], dtype=np.int64)
        self.n = np.array([A.shape[1] for A in self.blocks], dtype=np.int64)
        self.lda = np.array([_ld(A) for A in self.blocks], dtype=np.int64)
        a, b = opts.sketchfact_randn_samp
        adap = opts.sketchfact_adap or opts.rank < 0
        target = opts.nb if adap else opts.rank
        cap = max(1, a * target + b)
        while adap and a * 2 * target + b <= 72:   # the later stages' sketch orders
            target *= 2
            cap = a * target + b
        cap = int(ldt) if ldt is not None else cap
        self.ldt = np.array([max(1, min(int(n), cap)) for n in self.n], dtype=np.int64)
        self.sk = [torch.empty(max(int(n), 1), dtype=torch.int64, device="cuda") for n in self.n]
        self.rd = [torch.empty(max(int(n), 1), dtype=torch.int64, device="cuda") for n in self.n]
        self.T = [torch.empty(int(l) * max(int(n), 1), dtype=torch.float64, device="cuda")
                  for l, n in zip(self.ldt, self.n)]
        arr = C.c_void_p * max(nb, 1)
        self._A = arr(*[_ptr(A) for A in self.blocks])
        self._sk = arr(*[_ptr(t) for t in self.sk])
        self._rd = arr(*[_ptr(t) for t in self.rd])
        self._T = arr(*[_ptr(t) for t in self.T])
        self._copts = opts.to_c()
        self.k = np.full(nb, -1, dtype=np.int64)
        self.n_fallback = 0
        self.h = C.c_void_p()
        _lib.register_owner(self)
        check(_lib.load().lra_idfact_batched_plan(
            self.ctx.h, nb, self.m.ctypes.data, self.n.ctypes.data, C.cast(self._A, C.c_void_p),
            self.lda.ctypes.data, C.byref(self._copts), C.cast(self._sk, C.c_void_p),
            C.cast(self._rd, C.c_void_p), C.cast(self._T, C.c_void_p), self.ldt.ctypes.data,
            C.byref(self.h)), "lra_idfact_batched_plan")

    @property
    def widths(self) -> list:
        """Per adaptive round, the widest block run by a 512-thread workgroup (0: none); the
        plan's autotuner picks them on the blocks (they change time, not results)."""
        w = (C.c_int * 4)()
        check(_lib.load().lra_idfact_batched_plan_widths(self.h, w, 4),
              "lra_idfact_batched_plan_widths")
        return [int(x) for x in w if x >= 0]

    def capture(self) -> None:
        """Capture the launch in a hipGraph (needs a non-default current stream)."""
        check(_lib.load().lra_idfact_batched_capture(self.h), "lra_idfact_batched_capture")

    def launch(self) -> None:
        check(_lib.load().lra_idfact_batched_launch(self.h), "lra_idfact_batched_launch")

    def finish(self) -> list:
        i64v = C.c_int64(0)
        st = _lib.load().lra_idfact_batched_finish(self.h, self.k.ctypes.data, C.byref(i64v))
        over = []
        if st != 0:
            over = [b for b in range(self.nblk) if self.k[b] > self.ldt[b]]
            if st != _lib.E_UNSUPPORTED or not over:
                check(st, "lra_idfact_batched_finish")
        self.n_fallback = int(i64v.value)
        out = []
        for b in range(self.nblk):
            if b in over:  # rank beyond the reserved T rows: idfact with its own T
                out.append(idfact(self.blocks[b], self.opts))
                continue
            k, n, ldt = int(self.k[b]), int(self.n[b]), int(self.ldt[b])
            T = self.T[b][: ldt * max(n - k, 0)].reshape(max(n - k, 0), ldt).t()[:k]
            out.append(IDPackedV(self.sk[b][:k], self.rd[b][: n - k], T))
        return out

    def close(self) -> None:
        if self.h:
            _lib.load().lra_idfact_batched_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def idfact_batched(blocks, opts: LRAOptions = None, **kw) -> list:
    """idfact on every block of `blocks` (list of m_b x n_b matrices) through one persistent
    fused launch (MI355X extension): the same IDPackedV as idfact(A_b, opts) for each."""
    plan = BatchedID(blocks, opts, **kw)
    try:
        plan.launch()
        return plan.finish()
    finally:
        plan.close()


def _borrow_T(out, k: int, nr: int):
    """Copy the library-owned k x nr T (ldt = out.ldt) into a new column-major tensor."""
    torch = _torch()
    T = empty_colmajor(k, nr)
    if k > 0 and nr > 0:
        ldt = int(out.ldt)
        _copy_device(T.data_ptr(), int(out.T), k, nr, ldt, k)
    return T


def _copy_device(dst: int, src: int, rows: int, cols: int, lds: int, ldd: int) -> None:
    """2-D device copy of doubles on the current stream (hipMemcpy2DAsync via torch)."""
    torch = _torch()
    import ctypes
    hip = _hip()
    st = torch.cuda.current_stream().cuda_stream
    rc = hip.hipMemcpy2DAsync(ctypes.c_void_p(dst), ctypes.c_size_t(8 * ldd), ctypes.c_void_p(src),
                              ctypes.c_size_t(8 * lds), ctypes.c_size_t(8 * rows),
                              ctypes.c_size_t(cols), ctypes.c_int(3), ctypes.c_void_p(st))
    if rc != 0:
        hip.hipGetLastError()  # do not leave the error for the next launch check
        raise LraError(f"hipMemcpy2DAsync failed ({rc})")


_HIP = None


def _hip():
    global _HIP
    if _HIP is None:
        import ctypes
        _HIP = ctypes.CDLL("libamdhip64.so")
        _HIP.hipMemcpy2DAsync.restype = ctypes.c_int
    return _HIP


def id(trans, A=None, opts: LRAOptions = None, **kw):
    """id (src/id.jl:452-455): (sk, rd, T)."""
    V = idfact(trans, A, opts, **kw)
    return V.sk, V.rd, V.T


class LinearOperator:
    """Matrix-free operator (LinearOperator, src/linop.jl:9-40): `mul(X) -> L*X` and
    `mulc(X) -> L'*X` on float64 CUDA tensors (X is n x nrhs, resp. m x nrhs).  The library
    calls them through the C callback `lra_apply_fn` on torch's current stream."""

    def __init__(self, m: int, n: int, mul, mulc=None, hermitian: bool = False):
        self.m, self.n = m, n
        self.mul = mul
        self.mulc = mulc if mulc is not None else mul
        self.hermitian = hermitian

    @classmethod
    def from_matrix(cls, A, hermitian: bool = None):
        """LinOp(A) (src/linop.jl:42-60) of a dense matrix."""
        A = to_device(A)
        if hermitian is None:
            hermitian = _ishermitian_dev(A)
        return cls(A.shape[0], A.shape[1], lambda X: A @ X, lambda X: A.t() @ X, hermitian)

    @property
    def shape(self):
        return (self.m, self.n)

    def _callback(self):
        """lra_apply_fn: Y (ldy) = op(L) X (ldx); raw device pointers are copied into and out
        of torch tensors (a failing operator returns a non-zero status, never raises)."""
        def cb(user, trans, nrhs, X, ldx, Y, ldy, stream):
            try:
                rin = self.n if trans == 0 else self.m
                rout = self.m if trans == 0 else self.n
                Xt = empty_colmajor(rin, nrhs)
                if rin > 0 and nrhs > 0:
                    _copy_device(Xt.data_ptr(), X, rin, nrhs, ldx, rin)
                Yt = self.mul(Xt) if trans == 0 else self.mulc(Xt)
                Yt = to_device(Yt)
                if tuple(Yt.shape) != (rout, nrhs):
                    return 2
                if rout > 0 and nrhs > 0:
                    _copy_device(Y, Yt.data_ptr(), rout, nrhs, _ld(Yt), ldy)
                return 0
            except Exception as e:  # noqa: BLE001 -- reported to the library as a status
                self._error = e
                return 1
        self._error = None
        return _lib.APPLY_FN(cb)

    def _check(self, status: int, what: str) -> None:
        if status != 0 and getattr(self, "_error", None) is not None:
            raise LraError(f"{what}: the operator raised {self._error!r}") from self._error
        check(status, what)

    def _opts(self, opts, kw) -> "LRAOptions":
        """copy(opts; kw...) then chkopts!(opts, L) (src/LowRankApprox.jl:142-148): a sketch
        other than randn is replaced by randn with a warning."""
        opts = (opts or LRAOptions()).copy(**kw)
        if opts.sketch != "randn":
            warnings.warn('invalid sketch method; using "randn"', stacklevel=3)
            opts.sketch = "randn"
        return opts

    def idfact(self, trans="n", opts: "LRAOptions" = None, **kw) -> "IDPackedV":
        """idfact on the operator (src/id.jl:434-447 through src/linop.jl:105-113), with
        sketch_randn_niter power iterations through L and L' (src/sketch.jl:129-174)."""
        torch = _torch()
        opts = self._opts(opts, kw)
        opts.pqrfact_retval = "t"
        ctx = context()
        N = self.n if trans == "n" else self.m
        sk = torch.empty(max(N, 1), dtype=torch.int64, device="cuda")
        rd = torch.empty(max(N, 1), dtype=torch.int64, device="cuda")
        out = _lib.lra_id_out()
        out.sk, out.rd = _ptr(sk), _ptr(rd)
        out.T, out.ldt = None, 0
        copts = opts.to_c()
        cb = self._callback()
        self._check(_lib.load().lra_idfact_op(ctx.h, _lib.TRANS[trans], self.m, self.n, cb,
                                              None, int(self.hermitian), C.byref(copts),
                                              C.byref(out)), "lra_idfact_op")
        k = int(out.k)
        return IDPackedV(sk[:k], rd[: N - k], _borrow_T(out, k, N - k), int(out.rounds))

    def pqrfact(self, trans="n", opts: "LRAOptions" = None, **kw):
        """pqrfact on the operator (src/pqr.jl:285-320): skeleton columns op(L)[:, sk] from
        one application to one-hot vectors (src/util.jl:30-34, src/linop.jl:54-67)."""
        opts = self._opts(opts, kw)
        ctx = context()
        out = _lib.lra_pqr_out()
        copts = opts.to_c()
        cb = self._callback()
        self._check(_lib.load().lra_pqrfact_op(ctx.h, _lib.TRANS[trans], self.m, self.n, cb,
                                               None, int(self.hermitian), C.byref(copts),
                                               C.byref(out)), "lra_pqrfact_op")
        N = self.n if trans == "n" else self.m
        P = self.m if trans == "n" else self.n
        return _pqr_result(out, opts.pqrfact_retval, P, N)

    def psvdfact(self, opts: "LRAOptions" = None, **kw) -> "PartialSVD":
        """psvdfact on the operator (src/psvd.jl:238-272)."""
        opts = self._opts(opts, kw)
        ctx = context()
        out = _lib.lra_psvd_out()
        copts = opts.to_c()
        cb = self._callback()
        self._check(_lib.load().lra_psvdfact_op(ctx.h, self.m, self.n, cb, None,
                                                int(self.hermitian), C.byref(copts),
                                                C.byref(out)), "lra_psvdfact_op")
        return _psvd_result(out, self.m, self.n)

    def psvdvals(self, opts: "LRAOptions" = None, **kw):
        """psvdvals on the operator (src/psvd.jl:274-290)."""
        torch = _torch()
        opts = self._opts(opts, kw)
        ctx = context()
        S = torch.empty(max(min(self.m, self.n), 1), dtype=torch.float64, device="cuda")
        k = _lib.i64(0)
        copts = opts.to_c()
        cb = self._callback()
        self._check(_lib.load().lra_psvdvals_op(ctx.h, self.m, self.n, cb, None,
                                                int(self.hermitian), C.byref(copts), _ptr(S),
                                                C.byref(k)), "lra_psvdvals_op")
        return S[:int(k.value)]

    def pheigfact(self, opts: "LRAOptions" = None, **kw) -> "PartialHermEigen":
        """pheigfact on a Hermitian operator (src/pheig.jl:276-297): L[sk, sk] from one
        application to one-hot vectors (src/linop.jl:70-74)."""
        if self.m != self.n or not self.hermitian:
            raise ValueError("matrix must be Hermitian")
        opts = self._opts(opts, kw)
        ctx = context()
        out = _lib.lra_pheig_out()
        copts = opts.to_c()
        cb = self._callback()
        self._check(_lib.load().lra_pheigfact_op(ctx.h, self.n, cb, None, C.byref(copts),
                                                 C.byref(out)), "lra_pheigfact_op")
        return _pheig_result(out, self.n)

    def pheigvals(self, opts: "LRAOptions" = None, **kw):
        """pheigvals on a Hermitian operator (src/pheig.jl:298-314)."""
        torch = _torch()
        if self.m != self.n or not self.hermitian:
            raise ValueError("matrix must be Hermitian")
        opts = self._opts(opts, kw)
        ctx = context()
        v = torch.empty(max(self.n, 1), dtype=torch.float64, device="cuda")
        k = _lib.i64(0)
        copts = opts.to_c()
        cb = self._callback()
        self._check(_lib.load().lra_pheigvals_op(ctx.h, self.n, cb, None, C.byref(copts),
                                                 _ptr(v), C.byref(k)), "lra_pheigvals_op")
        return v[:int(k.value)]

    def curfact(self, opts: "LRAOptions" = None, **kw) -> "CURPackedU":
        """curfact on the operator (src/cur.jl:527-564)."""
        torch = _torch()
        opts = self._opts(opts, kw)
        ctx = context()
        mn = max(min(self.m, self.n), 1)
        rows = torch.empty(mn, dtype=torch.int64, device="cuda")
        cols = torch.empty(mn, dtype=torch.int64, device="cuda")
        k = _lib.i64(0)
        copts = opts.to_c()
        cb = self._callback()
        self._check(_lib.load().lra_curfact_op(ctx.h, self.m, self.n, cb, None,
                                               int(self.hermitian), C.byref(copts),
                                               _ptr(rows), _ptr(cols), C.byref(k)),
                    "lra_curfact_op")
        k = int(k.value)
        return CURPackedU(rows[:k], cols[:k])

    def getcols(self, cols, trans: str = "n"):
        """op(L)[:, cols] (getcols, src/util.jl:30-34) by one application to one-hot
        vectors (src/linop.jl:54-67)."""
        torch = _torch()
        cols = torch.as_tensor(cols, dtype=torch.int64, device="cuda")
        R = self.n if trans == "n" else self.m
        S = empty_colmajor(R, int(cols.shape[0]))
        S.zero_()
        if cols.numel():
            S[cols - 1, torch.arange(int(cols.shape[0]), device="cuda")] = 1.0
        return self.mul(S) if trans == "n" else self.mulc(S)

    def snorm(self, opts: "LRAOptions" = None, x0=None, **kw) -> float:
        """snorm on the operator (src/snorm.jl:14-45)."""
        opts = (opts or LRAOptions()).copy(**kw)
        ctx = context()
        x = to_device(np.asarray(x0, dtype=np.float64)) if x0 is not None else None
        res = C.c_double(0.0)
        copts = opts.to_c()
        cb = self._callback()
        self._check(_lib.load().lra_snorm_op(ctx.h, self.m, self.n, cb, None,
                                             int(self.hermitian), C.byref(copts), _ptr(x),
                                             C.byref(res)), "lra_snorm_op")
        return float(res.value)


# ----------------------------------------------------------------------------------------
# partial factorizations built on the ID

def _own(ptr: int, rows: int, cols: int, ld: int):
    """Copy a library-owned column-major device matrix into a new tensor."""
    out = empty_colmajor(rows, cols)
    if rows > 0 and cols > 0:
        _copy_device(out.data_ptr(), int(ptr), rows, cols, ld, rows)
    return out


def _own_vec(ptr: int, n: int, dtype=None):
    torch = _torch()
    out = torch.empty(max(n, 1), dtype=torch.float64 if dtype is None else dtype, device="cuda")
    if n > 0:
        _copy_device(out.data_ptr(), int(ptr), n, 1, n, n)
    return out[:n]


@dataclasses.dataclass
class PartialQR:
    """PartialQR (src/pqr.jl:25-31): op(A)[:, p] ~ Q R."""
    Q: object
    R: object
    p: object

    @property
    def k(self) -> int:
        return int(self.R.shape[0])

    def apply(self, X, trans: str = "n"):
        """op(A) X ~ Q R P' X = Q (R X[p, :]) (trans "n"), or op(A)' X = P R' (Q' X) ("c")
        (mul!, src/pqr.jl:86-131)."""
        torch = _torch()
        X, vec = _in_mat(X)
        if trans == "n":
            Xp = X[self.p - 1, :].t().contiguous().t()
            return _out_mat(dgemm(self.Q, dgemm(self.R, Xp)), vec)
        Y = dgemm(self.R, dgemm(self.Q, X, transa="c"), transa="c")  # R' (Q' X), rows in p order
        out = torch.empty_like(Y)
        out[self.p - 1, :] = Y
        return _out_mat(out.t().contiguous().t(), vec)

    def rapply(self, X, trans: str = "n"):
        """X op(A) = ((X Q) R) P' (trans "n") or X op(A)' = ((X P) R') Q' ("c")
        (right multiplication, src/pqr.jl:133-158)."""
        torch = _torch()
        X = to_device(X)
        vec = X.dim() == 1
        if vec:
            X = X.reshape(1, -1)
        if trans == "n":
            Y = dgemm(dgemm(X, self.Q), self.R)  # columns in p order
            out = torch.empty_like(Y)
            out[:, self.p - 1] = Y
            Y = out.t().contiguous().t()
        else:
            Xp = X[:, self.p - 1].t().contiguous().t()
            Y = dgemm(dgemm(Xp, self.R, transb="c"), self.Q, transb="c")
        return Y[0, :] if vec else Y

    def matrix(self):
        """Matrix(A) = Q R P' (src/pqr.jl:56)."""
        torch = _torch()
        return self.rapply(torch.eye(self.Q.shape[0], dtype=torch.float64, device="cuda"))

    def ldiv(self, B):
        """A \\ B, the minimum-norm least-squares solve P R^+ Q' B (ldiv!, src/pqr.jl:165-175).
        The reference forms the normal equations (R R') \\ (Q' B) then R' times that; here R^+
        comes from the one-sided Jacobi SVD of R' (n x k) = U diag(s) V', so
        R^+ = U diag(1/s) V', which equals R' (R R')^-1 for the full-row-rank R that pqrfact
        returns without squaring its condition number."""
        torch = _torch()
        B, vec = _in_mat(B)
        U, s, V = _gesvj(self.R.t())
        Y = dgemm(U, _scale_rows(1.0 / s, dgemm(V, dgemm(self.Q, B, transa="c"), transa="c")))
        out = torch.empty_like(Y)
        out[self.p - 1, :] = Y
        return _out_mat(out.t().contiguous().t(), vec)


@dataclasses.dataclass
class PQRFactors:
    """PQRFactors (src/pqr.jl:33-40): Q, R, p, k and T (entries not requested are None)."""
    Q: object
    R: object
    p: object
    k: int
    T: object


def _pqr_result(out, rv: str, P: int, N: int):
    """PartialQR (retval "qr") or PQRFactors from a filled lra_pqr_out."""
    torch = _torch()
    k = int(out.k)
    Q = _own(out.Q, P, k, int(out.ldq)) if "q" in rv and k > 0 else None
    R = _own(out.R, k, N, int(out.ldr)) if "r" in rv and k > 0 else None
    if "q" in rv and k == 0: Q = empty_colmajor(P, 0)
    if "r" in rv and k == 0: R = empty_colmajor(0, N)
    p = _own_vec(out.p, N, torch.int64)
    T = _own(out.T, k, N - k, int(out.ldt)) if "t" in rv and k > 0 and N > k else (
        empty_colmajor(k, N - k) if "t" in rv else None)
    if "q" in rv and "r" in rv and "t" not in rv:
        return PartialQR(Q, R, p)
    return PQRFactors(Q, R, p, k, T)


def pqrfact(trans, A=None, opts: LRAOptions = None, **kw):
    """pqrfact(trans, A, opts) (src/pqr.jl:285-320); PartialQR when pqrfact_retval is "qr"."""
    if A is None or isinstance(A, LRAOptions):
        trans, A, opts = "n", trans, A
    if isinstance(A, LinearOperator):
        return A.pqrfact(trans, opts, **kw)
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    N = n if trans == "n" else m
    P = m if trans == "n" else n
    out = _lib.lra_pqr_out()
    copts = opts.to_c()
    check(_lib.load().lra_pqrfact(ctx.h, _lib.TRANS[trans], m, n, _ptr(A), _ld(A),
                                  C.byref(copts), C.byref(out)), "lra_pqrfact")
    return _pqr_result(out, opts.pqrfact_retval, P, N)


def prange(trans, A=None, opts: LRAOptions = None, **kw):
    """prange(trans, A, opts) (src/prange.jl:14-77): Q with orthonormal columns spanning the
    numerical range of op(A) (trans "n" / "c") or of A and A' together ("b", square A)."""
    if A is None or isinstance(A, LRAOptions):
        trans, A, opts = "n", trans, A
    if trans not in _lib.PRANGE_TRANS:
        raise ValueError("trans")
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    P = n if trans == "c" else m
    out = _lib.lra_pqr_out()
    copts = opts.to_c()
    check(_lib.load().lra_prange(ctx.h, _lib.PRANGE_TRANS[trans], m, n, _ptr(A), _ld(A),
                                 C.byref(copts), C.byref(out)), "lra_prange")
    k = int(out.k)
    return _own(out.Q, P, k, int(out.ldq)) if k > 0 else empty_colmajor(P, 0)


def pqr(trans, A=None, opts: LRAOptions = None, **kw):
    """pqr (src/pqr.jl:310-315): (Q, R, p) with pqrfact_retval = "qr"."""
    if A is None or isinstance(A, LRAOptions):
        trans, A, opts = "n", trans, A
    kw.setdefault("pqrfact_retval", "qr")
    F = pqrfact(trans, A, opts, **kw)
    return F.Q, F.R, F.p


@dataclasses.dataclass
class PartialSVD:
    """PartialSVD (src/psvd.jl:16-24): A ~ U diag(S) Vt."""
    U: object
    S: object
    Vt: object

    @property
    def V(self):
        return self.Vt.t()

    @property
    def k(self) -> int:
        return int(self.S.shape[0])

    def apply(self, X, trans: str = "n"):
        """A X = U (S (Vt X)) or A' X = V (S (U' X)) (mul!, src/psvd.jl:61-124)."""
        X, vec = _in_mat(X)
        if trans == "n":
            Y = dgemm(self.U, _scale_rows(self.S, dgemm(self.Vt, X)))
        else:
            Y = dgemm(self.Vt, _scale_rows(self.S, dgemm(self.U, X, transa="c")), transa="c")
        return _out_mat(Y, vec)

    def rapply(self, X, trans: str = "n"):
        """X A (trans "n") or X A' ("c") (mul!, src/psvd.jl:99-116)."""
        return _right_apply(self, X, trans)

    def matrix(self):
        """Matrix(A) = (U diag(S)) Vt (src/psvd.jl:38)."""
        torch = _torch()
        return self.apply(torch.eye(self.Vt.shape[1], dtype=torch.float64, device="cuda"))

    def ldiv(self, B):
        """A \\ B = V (diag(S)^-1 (U' B)), the pseudo-inverse product (ldiv!,
        src/psvd.jl:119-124)."""
        B, vec = _in_mat(B)
        return _out_mat(dgemm(self.Vt, _scale_rows(1.0 / self.S, dgemm(self.U, B, transa="c")),
                              transa="c"), vec)


def _psvd_result(out, m: int, n: int) -> PartialSVD:
    k = int(out.k)
    return PartialSVD(_own(out.U, m, k, int(out.ldu)) if k else empty_colmajor(m, 0),
                      _own_vec(out.S, k), _own(out.Vt, k, n, int(out.ldvt)) if k else
                      empty_colmajor(0, n))


def psvdfact(A, opts: LRAOptions = None, **kw) -> PartialSVD:
    """psvdfact(A, opts) (src/psvd.jl:238-272)."""
    if isinstance(A, LinearOperator):
        return A.psvdfact(opts, **kw)
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    out = _lib.lra_psvd_out()
    copts = opts.to_c()
    check(_lib.load().lra_psvdfact(ctx.h, m, n, _ptr(A), _ld(A), C.byref(copts), C.byref(out)),
          "lra_psvdfact")
    return _psvd_result(out, m, n)


def psvdvals(A, opts: LRAOptions = None, **kw):
    """psvdvals(A, opts) (src/psvd.jl:274-290)."""
    torch = _torch()
    if isinstance(A, LinearOperator):
        return A.psvdvals(opts, **kw)
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    S = torch.empty(max(min(m, n), 1), dtype=torch.float64, device="cuda")
    k = _lib.i64(0)
    copts = opts.to_c()
    check(_lib.load().lra_psvdvals(ctx.h, m, n, _ptr(A), _ld(A), C.byref(copts), _ptr(S),
                                   C.byref(k)), "lra_psvdvals")
    return S[:int(k.value)]


def psvd(A, opts: LRAOptions = None, **kw):
    """psvd (src/psvd.jl:296-299): (U, S, V)."""
    F = psvdfact(A, opts, **kw)
    return F.U, F.S, F.V


@dataclasses.dataclass
class PartialHermEigen:
    """PartialHermEigen (src/pheig.jl:16-22): A ~ vectors diag(values) vectors'."""
    values: object
    vectors: object

    def apply(self, X, trans: str = "n"):
        """A X = V (w (V' X)) (mul!, src/pheig.jl:54-100); A' = A for the real Hermitian
        factorization, so trans is accepted for symmetry with the other factors."""
        X, vec = _in_mat(X)
        return _out_mat(dgemm(self.vectors, _scale_rows(self.values, dgemm(self.vectors, X,
                                                                            transa="c"))), vec)

    def rapply(self, X, trans: str = "n"):
        """X A = X A' (mul!, src/pheig.jl:114-145)."""
        return _right_apply(self, X, trans)

    def matrix(self):
        """Matrix(A) = (vectors diag(values)) vectors' (src/pheig.jl:31)."""
        torch = _torch()
        return self.apply(torch.eye(self.vectors.shape[0], dtype=torch.float64, device="cuda"))

    def ldiv(self, B):
        """A \\ B = V (diag(w)^-1 (V' B)), the pseudo-inverse product (ldiv!,
        src/pheig.jl:147-152)."""
        B, vec = _in_mat(B)
        return _out_mat(dgemm(self.vectors, _scale_rows(1.0 / self.values,
                                                        dgemm(self.vectors, B, transa="c"))),
                        vec)


def _pheig_result(out, n: int) -> PartialHermEigen:
    k = int(out.k)
    return PartialHermEigen(_own_vec(out.values, k),
                            _own(out.vectors, n, k, int(out.ldv)) if k else empty_colmajor(n, 0))


def pheigfact(A, opts: LRAOptions = None, **kw) -> PartialHermEigen:
    """pheigfact(A, opts) (src/pheig.jl:276-297) for a symmetric A."""
    if isinstance(A, LinearOperator):
        return A.pheigfact(opts, **kw)
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    n = A.shape[0]
    if A.shape[0] != A.shape[1] or not _ishermitian_dev(A):
        raise ValueError("matrix must be Hermitian")
    out = _lib.lra_pheig_out()
    copts = opts.to_c()
    check(_lib.load().lra_pheigfact(ctx.h, n, _ptr(A), _ld(A), C.byref(copts), C.byref(out)),
          "lra_pheigfact")
    return _pheig_result(out, n)


def pheigvals(A, opts: LRAOptions = None, **kw):
    """pheigvals(A, opts) (src/pheig.jl:298-314)."""
    torch = _torch()
    if isinstance(A, LinearOperator):
        return A.pheigvals(opts, **kw)
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    n = A.shape[0]
    if A.shape[0] != A.shape[1] or not _ishermitian_dev(A):
        raise ValueError("matrix must be Hermitian")
    v = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
    k = _lib.i64(0)
    copts = opts.to_c()
    check(_lib.load().lra_pheigvals(ctx.h, n, _ptr(A), _ld(A), C.byref(copts), _ptr(v),
                                    C.byref(k)), "lra_pheigvals")
    return v[:int(k.value)]


def pheig(A, opts: LRAOptions = None, **kw):
    """pheig (src/pheig.jl:317-320): (values, vectors)."""
    F = pheigfact(A, opts, **kw)
    return F.values, F.vectors


def sketchfact(trans, A=None, opts: "LRAOptions" = None, stream_base: int = 0, **kw):
    """sketchfact(:left, trans, A, opts) with pqrfact_retval = "" (src/sketch.jl:52-66):
    (column pivots of op(A), 1-based, device; rank k)."""
    torch = _torch()
    if A is None or isinstance(A, LRAOptions):
        trans, A, opts = "n", trans, A
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    N = n if trans == "n" else m
    p = torch.empty(max(N, 1), dtype=torch.int64, device="cuda")
    k = _lib.i64(0)
    copts = opts.to_c()
    check(_lib.load().lra_sketchfact(ctx.h, _lib.TRANS[trans], m, n, _ptr(A), _ld(A),
                                     C.byref(copts), stream_base, _ptr(p), C.byref(k)),
          "lra_sketchfact")
    return p[:N], int(k.value)


@dataclasses.dataclass
class CURPackedU:
    """CURPackedU (src/cur.jl:16-24): row and column skeletons (1-based, device)."""
    rows: object
    cols: object

    def __getitem__(self, key):
        if key == "rows": return self.rows
        if key == "cols": return self.cols
        if key == "k": return int(self.rows.shape[0])
        raise KeyError(key)


@dataclasses.dataclass
class CUR:
    """CUR (src/cur.jl:60-66): A ~ C U R with C = A[:, cols], R = A[rows, :] and U the
    pseudo-inverse of the intersection A[rows, cols] as PartialSVD(V, 1 ./ sigma, U') (the
    construction at src/cur.jl:85-93; for a symmetric skeleton rows = cols, the HermCUR /
    SymCUR forms of :102-116 invert the same intersection)."""
    rows: object
    cols: object
    C: object
    U: PartialSVD
    R: object

    @classmethod
    def from_packed(cls, A, U: "CURPackedU") -> "CUR":
        """CUR(A, U::CURPackedU)."""
        return cls.from_skeletons(A, U.rows, U.cols)

    @classmethod
    def from_skeletons(cls, A, rows, cols) -> "CUR":
        """CUR(A, rows, cols) (1-based skeletons)."""
        torch = _torch()
        ctx = context()
        A = to_device(A)
        m, n = A.shape
        def idx(v):
            if not isinstance(v, torch.Tensor):
                v = torch.as_tensor(np.asarray(v, dtype=np.int64))
            return v.to(device="cuda", dtype=torch.int64).contiguous()
        rows, cols = idx(rows), idx(cols)
        k = int(rows.shape[0])
        if int(cols.shape[0]) != k:
            raise ValueError("rows and cols differ in length")
        Cm, Rm = empty_colmajor(m, k), empty_colmajor(k, n)
        V, Ut = empty_colmajor(k, k), empty_colmajor(k, k)
        sinv = torch.empty(k, dtype=torch.float64, device="cuda")
        check(_lib.load().lra_cur_build(ctx.h, m, n, _ptr(A), _ld(A), _ptr(rows), _ptr(cols), k,
                                        _ptr(Cm), max(1, m), _ptr(Rm), max(1, k), _ptr(V),
                                        max(1, k), _ptr(sinv), _ptr(Ut), max(1, k)),
              "lra_cur_build")
        return cls(rows, cols, Cm, PartialSVD(V, sinv, Ut), Rm)

    def apply(self, X, trans: str = "n"):
        """C (U (R X)) (trans "n") or R' (U' (C' X)) ("c") (mul!, src/cur.jl:209-523)."""
        X, vec = _in_mat(X)
        if trans == "n":
            Y = dgemm(self.C, self.U.apply(dgemm(self.R, X)))
        else:
            Y = dgemm(self.R, self.U.apply(dgemm(self.C, X, transa="c"), "c"), transa="c")
        return _out_mat(Y, vec)

    def rapply(self, X, trans: str = "n"):
        """X A (trans "n") or X A' ("c") (*, src/cur.jl:462-523)."""
        return _right_apply(self, X, trans)

    def matrix(self):
        """Matrix(A) = C (U R) (src/cur.jl:153)."""
        torch = _torch()
        return self.apply(torch.eye(self.R.shape[1], dtype=torch.float64, device="cuda"))


def curfact(A, opts: "LRAOptions" = None, **kw) -> CURPackedU:
    """curfact(A, opts) (src/cur.jl:532-564)."""
    torch = _torch()
    if isinstance(A, LinearOperator):
        return A.curfact(opts, **kw)
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    mn = max(min(m, n), 1)
    rows = torch.empty(mn, dtype=torch.int64, device="cuda")
    cols = torch.empty(mn, dtype=torch.int64, device="cuda")
    k = _lib.i64(0)
    copts = opts.to_c()
    check(_lib.load().lra_curfact(ctx.h, m, n, _ptr(A), _ld(A), C.byref(copts), _ptr(rows),
                                  _ptr(cols), C.byref(k)), "lra_curfact")
    k = int(k.value)
    return CURPackedU(rows[:k], cols[:k])


def cur(A, opts: "LRAOptions" = None, **kw):
    """cur (src/cur.jl:566-569): (rows, cols)."""
    U = curfact(A, opts, **kw)
    return U.rows, U.cols


# ----------------------------------------------------------------------------------------
# snorm / snormdiff (src/snorm.jl:14-53)

def _ishermitian_dev(A) -> bool:
    torch = _torch()
    return A.shape[0] == A.shape[1] and bool(torch.equal(A, A.t()))


def snorm(A, opts: LRAOptions = None, x0=None, **kw) -> float:
    """Spectral norm by power iteration on A'A (A itself when Hermitian)."""
    if isinstance(A, LinearOperator):
        return A.snorm(opts, x0, **kw)
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    x = to_device(np.asarray(x0, dtype=np.float64)) if x0 is not None else None
    res = C.c_double(0.0)
    copts = opts.to_c()
    check(_lib.load().lra_snorm(ctx.h, m, n, _ptr(A), _ld(A), int(_ishermitian_dev(A)),
                                C.byref(copts), _ptr(x), C.byref(res)), "lra_snorm")
    return float(res.value)


def snormdiff(A, B, opts: LRAOptions = None, x0=None, **kw) -> float:
    """snorm(A - B); B may be a dense matrix or an IDPackedV of A (applied matrix-free)."""
    opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
    ctx = context()
    A = to_device(A)
    m, n = A.shape
    x = to_device(np.asarray(x0, dtype=np.float64)) if x0 is not None else None
    res = C.c_double(0.0)
    copts = opts.to_c()
    L = _lib.load()
    if isinstance(B, IDPackedV):
        T = to_device(B.T) if B.T.numel() else empty_colmajor(max(B.k, 1), 1)
        check(L.lra_snormdiff_id(ctx.h, m, n, _ptr(A), _ld(A), B.k, _ptr(B.sk), _ptr(B.rd),
                                 _ptr(T), max(1, _ld(T)), C.byref(copts), _ptr(x),
                                 C.byref(res)), "lra_snormdiff_id")
    else:
        Bd = to_device(B)
        herm = int(_ishermitian_dev(A) and _ishermitian_dev(Bd))
        check(L.lra_snormdiff(ctx.h, m, n, _ptr(A), _ld(A), _ptr(Bd), _ld(Bd), herm,
                              C.byref(copts), _ptr(x), C.byref(res)), "lra_snormdiff")
    return float(res.value)


# ----------------------------------------------------------------------------------------
# serving

class Server:
    """Serving loop for idfact / psvdfact requests on device-resident matrices (lra_server_*):
    requests are admitted in arrival order when a worker is free and their memory estimate
    fits `hbm_budget` next to the factors held; `workers` host threads, each with its own
    context and stream, overlap their requests' kernels; factors stay on the device until
    `release`.  Each result equals idfact / psvdfact of the request alone."""

    KINDS = {"idfact": 0, "psvdfact": 1}
    STATUS = {0: "queued", 1: "running", 2: "done", 3: "failed"}

    def __init__(self, hbm_budget: int = 64 << 30, workers: int = 4,
                 device: Optional[int] = None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise LraError("no GPU visible: the MI355X path has no CPU fallback")
        self.device = torch.cuda.current_device() if device is None else device
        self.h = C.c_void_p()
        check(_lib.load().lra_server_create(self.device, int(hbm_budget), int(workers),
                                            C.byref(self.h)), "lra_server_create")
        _lib.register_owner(self)
        self._keep = {}  # handle -> (A, kind): A stays alive while the request may read it

    def submit(self, A, kind: str = "idfact", opts: LRAOptions = None, **kw) -> int:
        torch = _torch()
        opts = (opts or LRAOptions.for_type(_eltype(A))).copy(**kw)
        if kind == "idfact":
            opts.pqrfact_retval = "t"
        A = to_device(A)
        # the workers' streams do not wait for torch's: A must be complete now
        torch.cuda.current_stream().synchronize()
        copts = opts.to_c()
        h = _lib.i64(0)
        check(_lib.load().lra_server_submit(self.h, self.KINDS[kind], A.shape[0], A.shape[1],
                                            _ptr(A), _ld(A), C.byref(copts), C.byref(h)),
              "lra_server_submit")
        self._keep[int(h.value)] = (A, kind)
        return int(h.value)

    def poll(self, h: int) -> str:
        st = C.c_int32(0)
        check(_lib.load().lra_server_poll(self.h, h, C.byref(st)), "lra_server_poll")
        return self.STATUS[st.value]

    def wait(self, h: int, timeout: float = -1.0) -> str:
        st = C.c_int32(0)
        check(_lib.load().lra_server_wait(self.h, h, float(timeout), C.byref(st)),
              "lra_server_wait")
        return self.STATUS[st.value]

    def _out(self, h: int):
        out = _lib.lra_server_out()
        check(_lib.load().lra_server_result(self.h, h, C.byref(out)), "lra_server_result")
        return out

    def times(self, h: int):
        """(submitted, started, done) in seconds on the server's clock."""
        out = self._out(h)
        return out.t_submit, out.t_start, out.t_done

    def result(self, h: int):
        """Copy of the factors: IDPackedV (idfact) or PartialSVD (psvdfact)."""
        torch = _torch()
        out = self._out(h)
        m, n, k = int(out.m), int(out.n), int(out.k)
        hip = _hip()
        st = torch.cuda.current_stream().cuda_stream
        if out.kind == 0:
            sk = torch.empty(max(k, 1), dtype=torch.int64, device="cuda")
            rd = torch.empty(max(n - k, 1), dtype=torch.int64, device="cuda")
            if k > 0:
                hip.hipMemcpyAsync(C.c_void_p(sk.data_ptr()), C.c_void_p(out.sk),
                                   C.c_size_t(8 * k), C.c_int(3), C.c_void_p(st))
            if n - k > 0:
                hip.hipMemcpyAsync(C.c_void_p(rd.data_ptr()), C.c_void_p(out.rd),
                                   C.c_size_t(8 * (n - k)), C.c_int(3), C.c_void_p(st))
            T = empty_colmajor(k, n - k)
            if k > 0 and n - k > 0:
                _copy_device(T.data_ptr(), int(out.T), k, n - k, int(out.ldt), k)
            return IDPackedV(sk[:k], rd[: n - k], T)
        U = empty_colmajor(m, k)
        Vt = empty_colmajor(k, n)
        S = torch.empty(k, dtype=torch.float64, device="cuda")
        if k > 0:
            _copy_device(U.data_ptr(), int(out.U), m, k, int(out.ldu), m)
            _copy_device(Vt.data_ptr(), int(out.Vt), k, n, int(out.ldvt), k)
            _copy_device(S.data_ptr(), int(out.S), k, 1, k, k)
        return PartialSVD(U, S, Vt)

    def release(self, h: int) -> None:
        torch = _torch()
        torch.cuda.current_stream().synchronize()  # copies made by result() read the factors
        check(_lib.load().lra_server_release(self.h, h), "lra_server_release")
        self._keep.pop(h, None)

    def stats(self):
        q, r, b = _lib.i64(0), _lib.i64(0), C.c_size_t(0)
        check(_lib.load().lra_server_stats(self.h, C.byref(q), C.byref(r), C.byref(b)),
              "lra_server_stats")
        return {"queued": int(q.value), "running": int(r.value), "bytes_held": int(b.value)}

    def close(self) -> None:
        if self.h:
            _lib.load().lra_server_destroy(self.h)
            self.h = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
