"""CPU parity oracle for the sketch-then-partial-factorize path of LowRankApprox.jl.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and bench.py's
CPU-baseline leg import this module, and only as the checker / the timed CPU baseline.  The
product path (``lowrankapprox.jl_amd`` + ``liblra_hip.so``) never imports it.

It restates the reference's driver code in numpy on top of the C restatement in
``lra_oracle.c`` (QRCP block step, rank-adaptive GEQP3, T = R11 \\ R12, strong-RRQR swaps,
SRFT / subset / sparse-Gaussian sketches).  LAPACK calls the reference makes through Julia's
LinearAlgebra (qr!, svd!, eigen!) are made here through numpy/scipy, i.e. the same LAPACK
routines.  All randomness is injected through a ``Rand`` provider so the GPU path and the
oracle see identical sketch matrices.

Parity pins (tests/test_oracle.py): the restated dlaqps against OpenBLAS's dlaqps (the
routine the reference ccall's at src/lapack.jl:129-136), geqp3_adap at full rank against
scipy's dgeqp3, Hilbert ranks against the reference README (26/27 at the default rtol, ~22 at
1e-12, 20 with rank=20), the SRFT against numpy.fft, snorm against numpy.linalg.norm(., 2),
psvd singular values against matrices built with known singular values.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import glob
import os
import subprocess
from typing import Callable, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_OPENBLAS = None

i64 = C.c_int64
P_D = C.POINTER(C.c_double)
P_I = C.POINTER(C.c_int64)


def build(force: bool = False) -> str:
    """Compile liblra_oracle.so with the committed Makefile (gcc)."""
    so = os.path.join(_HERE, "liblra_oracle.so")
    src = os.path.join(_HERE, "lra_oracle.c")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liblra_oracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        L.orc_geqp3_adap.restype = i64
        L.orc_geqp3_adap.argtypes = [i64, i64, P_D, i64, i64, C.c_double, C.c_double, i64,
                                     P_I, P_D, C.c_void_p, P_D]
        L.orc_dlaqps_gap.restype = None
        L.orc_dlaqps_gap.argtypes = [i64, i64, i64, i64, P_I, P_D, i64, P_I, P_D, P_D, P_D,
                                     P_D, P_D, i64, P_D]
        L.orc_trsm_upper.restype = None
        L.orc_trsm_upper.argtypes = [i64, i64, P_D, i64, P_D, i64]
        L.orc_maxdet_swapcols.restype = i64
        L.orc_maxdet_swapcols.argtypes = [i64, i64, P_I, P_D, i64, P_D, i64, C.c_double, i64]
        L.orc_srft_l.restype = i64
        L.orc_srft_l.argtypes = [i64, i64]
        L.orc_srft_left.restype = None
        L.orc_srft_left.argtypes = [i64, i64, P_D, i64, i64, P_D, P_I, P_D, i64]
        L.orc_subset_left.restype = None
        L.orc_subset_left.argtypes = [i64, i64, P_D, i64, i64, P_I, P_D, i64]
        L.orc_sprn_left.restype = None
        L.orc_sprn_left.argtypes = [i64, i64, P_D, i64, i64, P_I, P_D, P_D, i64]
        _LIB = L
    return _LIB


def openblas():
    """The ILP64 OpenBLAS bundled with numpy (called only, never inspected)."""
    global _OPENBLAS
    if _OPENBLAS is None:
        d = os.path.join(os.path.dirname(np.__file__), "..", "numpy.libs")
        cands = sorted(glob.glob(os.path.join(d, "libscipy_openblas64_*.so")))
        if not cands:
            raise RuntimeError("numpy's OpenBLAS (ILP64) not found")
        _OPENBLAS = C.CDLL(cands[0])
    return _OPENBLAS


def openblas_laqps_ptr() -> int:
    return C.cast(openblas().scipy_dlaqps_64_, C.c_void_p).value


def openblas_set_threads(n: int) -> None:
    openblas().scipy_openblas_set_num_threads64_(C.c_int(int(n)))


def _pd(a):
    return a.ctypes.data_as(P_D)


def _pi(a):
    return a.ctypes.data_as(P_I)


# ----------------------------------------------------------------------------------------
# options (src/LowRankApprox.jl:77-148)

@dataclasses.dataclass
class LRAOptions:
    """Mirror of LRAOptions with the reference defaults for Float64
    (src/LowRankApprox.jl:96-119).  The samp functions are affine n -> a*n + b."""
    atol: float = 0.0
    maxdet_niter: int = -1
    maxdet_tol: float = -1.0
    nb: int = 32
    pheig_orthtol: float = float(np.sqrt(np.finfo(np.float64).eps))
    pqrfact_retval: str = "qr"
    rank: int = -1
    rtol: float = 5 * float(np.finfo(np.float64).eps)
    sketch: str = "randn"
    sketch_randn_niter: int = 0
    sketchfact_adap: bool = True
    sketchfact_randn_samp: tuple = (1, 8)
    sketchfact_srft_samp: tuple = (1, 8)
    sketchfact_sub_samp: tuple = (4, 8)
    snorm_niter: int = 32
    verb: bool = True

    def copy(self, **kw) -> "LRAOptions":
        o = dataclasses.replace(self)
        for k, v in kw.items():
            if not hasattr(o, k):
                raise AttributeError(k)
            setattr(o, k, v)
        return o


def chkopts(opts: LRAOptions) -> None:
    """chkopts! (src/LowRankApprox.jl:133-141)."""
    if not opts.atol >= 0: raise ValueError("atol")
    if not opts.nb > 0: raise ValueError("nb")
    if not opts.pheig_orthtol >= 0: raise ValueError("pheig_orthtol")
    if not opts.rtol >= 0: raise ValueError("rtol")
    if opts.sketch not in ("none", "randn", "sprn", "srft", "sub"): raise ValueError("sketch")
    opts.pqrfact_retval = opts.pqrfact_retval.lower()


def samp(fn: tuple, n: int) -> int:
    return fn[0] * n + fn[1]


# ----------------------------------------------------------------------------------------
# injected randomness

class Rand:
    """Source of the random objects of each adaptive round.  ``gaussian(round, rows, cols)``
    returns the rows x cols Gaussian sketch; ``srft(round, m, k)`` returns (d, idx);
    ``subset(round, m, k)`` the 1-based sample rows; ``sprn(round, m, k)`` (perm, coeffs);
    ``start(n)`` the power-iteration start vector.  The default draws from numpy."""

    def __init__(self, seed: int = 0):
        self.rng = np.random.default_rng(seed)

    def gaussian(self, rnd: int, rows: int, cols: int) -> np.ndarray:
        return np.asfortranarray(self.rng.standard_normal((rows, cols)))

    def srft(self, rnd: int, m: int, k: int):
        d = np.where(self.rng.random(m) > 0.5, 1.0, -1.0)
        idx = self.rng.integers(1, m + 1, size=k).astype(np.int64)
        return d, idx

    def subset(self, rnd: int, m: int, k: int):
        return self.rng.integers(1, m + 1, size=k).astype(np.int64)

    def sprn(self, rnd: int, m: int, k: int):
        return (self.rng.permutation(m) + 1).astype(np.int64), self.rng.standard_normal(m)

    def start(self, n: int) -> np.ndarray:
        return self.rng.standard_normal(n)


# ----------------------------------------------------------------------------------------
# C-backed kernels

def geqp3_adap(B: np.ndarray, rank: int, atol: float, rtol: float, nb: int,
               use_openblas: bool = False, want_gap: bool = False):
    """geqp3_adap! (src/pqr.jl:348-418) on a copy of B.  Returns (Bfact, jpvt 1-based, tau,
    k[, gap])."""
    B = np.array(B, dtype=np.float64, order="F", copy=True)
    m, n = B.shape
    jpvt = np.zeros(n, dtype=np.int64)
    tau = np.zeros(max(1, min(m, n)), dtype=np.float64)
    gap = np.ones(max(1, min(m, n)), dtype=np.float64)
    fn = openblas_laqps_ptr() if use_openblas else None
    maxnrm = float(np.linalg.norm(B, axis=0).max()) if B.size else 0.0
    k = lib().orc_geqp3_adap(m, n, _pd(B), max(1, m), rank, atol, rtol, nb, _pi(jpvt),
                             _pd(tau), fn, _pd(gap) if (want_gap and not use_openblas) else None)
    if want_gap:
        # norm gap at each pivot step relative to the largest initial column norm
        return B, jpvt, tau, int(k), gap / maxnrm if maxnrm > 0 else gap
    return B, jpvt, tau, int(k)


def laqps(A: np.ndarray, offset: int, nb: int, jpvt, vn1, vn2, use_openblas=False):
    """One dlaqps block step (restated or OpenBLAS).  Returns (A, jpvt, tau, vn1, vn2, kb)."""
    A = np.array(A, dtype=np.float64, order="F", copy=True)
    m, n = A.shape
    jpvt = np.array(jpvt, dtype=np.int64, copy=True)
    vn1 = np.array(vn1, dtype=np.float64, copy=True)
    vn2 = np.array(vn2, dtype=np.float64, copy=True)
    tau = np.zeros(max(1, nb), dtype=np.float64)
    auxv = np.zeros(max(1, nb), dtype=np.float64)
    F = np.zeros((max(1, n), max(1, nb)), dtype=np.float64, order="F")
    kb = np.zeros(1, dtype=np.int64)
    if use_openblas:
        fn = openblas().scipy_dlaqps_64_
        fn.restype = None
        a = [C.byref(i64(m)), C.byref(i64(n)), C.byref(i64(offset)), C.byref(i64(nb)), _pi(kb),
             _pd(A), C.byref(i64(max(1, m))), _pi(jpvt), _pd(tau), _pd(vn1), _pd(vn2),
             _pd(auxv), _pd(F), C.byref(i64(max(1, n)))]
        fn(*a)
    else:
        lib().orc_dlaqps_gap(m, n, offset, nb, _pi(kb), _pd(A), max(1, m), _pi(jpvt), _pd(tau),
                             _pd(vn1), _pd(vn2), _pd(auxv), _pd(F), max(1, n), None)
    return A, jpvt, tau[: int(kb[0])], vn1, vn2, int(kb[0])


def maxdet_t(R: np.ndarray) -> np.ndarray:
    """T = R[:,1:k] \\ R[:,k+1:n] (src/pqr.jl:438-442)."""
    R = np.asfortranarray(R, dtype=np.float64)
    k, n = R.shape
    T = np.zeros((k, max(0, n - k)), dtype=np.float64, order="F")
    if k > 0 and n > k:
        lib().orc_trsm_upper(k, n, _pd(R), max(1, k), _pd(T), max(1, k))
    return T


# ----------------------------------------------------------------------------------------
# complex element type (T <: Complex): zlaqps through the reference's loop

def zlaqps(A: np.ndarray, offset: int, nb: int, jpvt, vn1, vn2):
    """One OpenBLAS zlaqps block (the ccall at src/lapack.jl:117-139 for ComplexF64; called
    only) on A in place; returns kb and tau (kb entries)."""
    ob = openblas()
    m, n = A.shape
    kb = np.zeros(1, dtype=np.int64)
    tau = np.zeros(max(1, nb), dtype=np.complex128)
    auxv = np.zeros(max(1, nb), dtype=np.complex128)
    F = np.zeros((max(1, n), max(1, nb)), dtype=np.complex128, order="F")
    pz = lambda a: a.ctypes.data_as(C.c_void_p)
    ob.scipy_zlaqps_64_(C.byref(i64(m)), C.byref(i64(n)), C.byref(i64(offset)),
                        C.byref(i64(nb)), pz(kb), pz(A), C.byref(i64(max(1, m))), pz(jpvt),
                        pz(tau), pz(vn1), pz(vn2), pz(auxv), pz(F), C.byref(i64(max(1, n))))
    return int(kb[0]), tau[: int(kb[0])]


def zgeqp3_adap(B: np.ndarray, rank: int, atol: float, rtol: float, nb: int):
    """geqp3_adap! + geqp3_adap_main! (src/pqr.jl:348-418) for a complex B, restated over
    OpenBLAS zlaqps exactly as the reference's complex branch calls it (rwork norms, blocks
    of min(nb, k-j+1) columns with offset j-1, the |A(jn-1,jn-1)| <= ptol test after each
    block).  Returns (Bfact, jpvt 1-based, tau, k)."""
    A = np.array(B, dtype=np.complex128, order="F", copy=True)
    m, n = A.shape
    jpvt = np.arange(1, n + 1, dtype=np.int64)
    l = min(m, n)
    k = l if (rank < 0 or rank > l) else rank
    tau = np.zeros(max(1, k), dtype=np.complex128)
    if k <= 0:
        return A, jpvt, tau[:0], k
    nbk = min(nb, k)
    rwork = np.zeros(2 * n)
    for j in range(n):
        rwork[j] = rwork[n + j] = np.linalg.norm(A[:, j])
    ptol = max(atol, rtol * rwork[:n].max())
    j = 1
    while j <= k:
        jb = min(nbk, k - j + 1)
        sub = np.asfortranarray(A[:, j - 1:])
        jp = np.ascontiguousarray(jpvt[j - 1:])
        v1 = np.ascontiguousarray(rwork[j - 1:n])
        v2 = np.ascontiguousarray(rwork[n + j - 1:2 * n])
        fjb, t = zlaqps(sub, j - 1, jb, jp, v1, v2)
        A[:, j - 1:] = sub
        jpvt[j - 1:] = jp
        rwork[j - 1:n] = v1
        rwork[n + j - 1:2 * n] = v2
        tau[j - 1:j - 1 + fjb] = t
        jn = j + fjb
        if abs(A[jn - 2, jn - 2]) <= ptol:
            for i in range(j, jn):
                if abs(A[i - 1, i - 1]) <= ptol:
                    return A, jpvt, tau[:i - 1], i - 1
        j = jn
    return A, jpvt, tau[:k], k


def zsketch_phi(sketch: str, m: int, order: int, draws) -> np.ndarray:
    """The dense complex left-sketch matrix (order x m) of each kind, from injected draws:
    randn: crandn (src/util.jl:4-11) given as the matrix itself; srft: Phi[p, t] =
    d_t exp(-2 pi i f_p t / m), f_p = idx_p - 1 -- the sampled DFT the reference's six-step
    complex SRFT computes (src/sketch.jl:366-376, 452-472); sub: one-hot rows (src/sketch.jl:
    248-257); sprn: row p holds s at columns r of its group (src/sketch.jl:571-589)."""
    if sketch == "randn":
        return np.asarray(draws, dtype=np.complex128)
    Phi = np.zeros((order, m), dtype=np.complex128)
    if sketch == "srft":
        d, idx = draws
        f = np.asarray(idx, dtype=np.int64) - 1
        t = np.arange(m, dtype=np.int64)
        e = (f[:, None] * t[None, :]) % m
        Phi = np.exp(-2j * np.pi * e / m) * np.asarray(d)[None, :]
        return Phi
    if sketch == "sub":
        r = np.asarray(draws, dtype=np.int64)
        Phi[np.arange(order), r - 1] = 1.0
        return Phi
    if sketch == "sprn":
        r, sc = draws
        o = 0
        for p in range(order):
            q = (m - p - 1) // order + 1 if m - p - 1 >= 0 else 0
            Phi[p, np.asarray(r[o:o + q]) - 1] = sc[o:o + q]
            o += q
        return Phi
    raise ValueError(sketch)


def zidfact(trans: str, A, opts: LRAOptions = None, zrand=None, **kw):
    """idfact (src/id.jl:434-447) for a complex A: the adaptive sketch loop
    (src/sketch.jl:223-240 and siblings) with injected complex sketches (zrand.phi(sketch,
    round, m, order) -> dense order x m matrix), geqp3_adap over OpenBLAS zlaqps and
    T = R11 \\ R12 (scipy's solve_triangular, LAPACK ztrtrs).  sketch = none factors op(A)."""
    import scipy.linalg as sla
    opts = (opts or LRAOptions()).copy(**kw)
    chkopts(opts)
    A = np.asarray(A, dtype=np.complex128)
    X = A if trans == "n" else A.conj().T
    P, N = X.shape
    rounds = 1
    if opts.sketch == "none":
        Bf, p, tau, k = zgeqp3_adap(X, opts.rank, opts.atol, opts.rtol, opts.nb)
    else:
        sampfn = {"randn": opts.sketchfact_randn_samp, "srft": opts.sketchfact_srft_samp,
                  "sub": opts.sketchfact_sub_samp, "sprn": (1, 0)}[opts.sketch]
        adap = opts.sketchfact_adap or opts.rank < 0
        target = opts.nb if adap else opts.rank
        rnd = 0
        while True:
            order = samp(sampfn, target)
            B = zrand.phi(opts.sketch, rnd, P, order) @ X
            Bf, p, tau, k = zgeqp3_adap(B, opts.rank, opts.atol, opts.rtol, opts.nb)
            rounds = rnd + 1
            if not adap or k < target:
                break
            target *= 2
            rnd += 1
    R = np.triu(Bf[:k, :])
    T = sla.solve_triangular(R[:, :k], R[:, k:]) if 0 < k < N else np.zeros((k, N - k),
                                                                            np.complex128)
    return IDPackedV(p[:k].copy(), p[k:].copy(), T, rounds, None)


def maxdet_t_blas(R: np.ndarray) -> np.ndarray:
    """T = R[:,1:k] \\ R[:,k+1:n] through OpenBLAS dtrsm (the BLAS call behind
    ldiv!(UpperTriangular(R1), R2) at src/pqr.jl:441; called only): the CPU-baseline form of
    maxdet_t."""
    R = np.asfortranarray(R, dtype=np.float64)
    k, n = R.shape
    T = np.array(R[:, k:], dtype=np.float64, order="F", copy=True)
    if k > 0 and n > k:
        ob = openblas()
        one = C.c_double(1.0)
        side, uplo, tr, diag = (C.c_char(x) for x in (b"L", b"U", b"N", b"N"))
        kk, nn, ld = i64(k), i64(n - k), i64(max(1, k))
        ob.scipy_dtrsm_64_(C.byref(side), C.byref(uplo), C.byref(tr), C.byref(diag),
                           C.byref(kk), C.byref(nn), C.byref(one), _pd(R), C.byref(ld), _pd(T),
                           C.byref(ld), i64(1), i64(1), i64(1), i64(1))
    return T


def maxdet_swapcols(p: np.ndarray, T: np.ndarray, tol: float, niter: int, R1=None):
    """maxdet_swapcols!/maxdet_update! (src/pqr.jl:445-501) on T/p (and R1 if given)."""
    T = np.array(T, dtype=np.float64, order="F", copy=True)
    p = np.array(p, dtype=np.int64, copy=True)
    k, nr = T.shape
    r1 = None if R1 is None else np.array(R1, dtype=np.float64, order="F", copy=True)
    it = lib().orc_maxdet_swapcols(k, k + nr, _pi(p), _pd(T), max(1, k),
                                   _pd(r1) if r1 is not None else None, max(1, k), tol, niter)
    return p, T, r1, int(it)


def srft_l(m: int, k: int) -> int:
    return int(lib().orc_srft_l(m, k))


def srft_left(A: np.ndarray, k: int, d: np.ndarray, idx: np.ndarray) -> np.ndarray:
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    d = np.ascontiguousarray(d, dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    Cm = np.zeros((k, n), dtype=np.float64, order="F")
    lib().orc_srft_left(m, n, _pd(A), max(1, m), k, _pd(d), _pi(idx), _pd(Cm), max(1, k))
    return Cm


def subset_left(A: np.ndarray, k: int, r: np.ndarray) -> np.ndarray:
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    r = np.ascontiguousarray(r, dtype=np.int64)
    Cm = np.zeros((k, n), dtype=np.float64, order="F")
    lib().orc_subset_left(m, n, _pd(A), max(1, m), k, _pi(r), _pd(Cm), max(1, k))
    return Cm


def sprn_left(A: np.ndarray, k: int, r: np.ndarray, s: np.ndarray) -> np.ndarray:
    A = np.asfortranarray(A, dtype=np.float64)
    m, n = A.shape
    r = np.ascontiguousarray(r, dtype=np.int64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    Cm = np.zeros((k, n), dtype=np.float64, order="F")
    lib().orc_sprn_left(m, n, _pd(A), max(1, m), k, _pi(r), _pd(s), _pd(Cm), max(1, k))
    return Cm


# ----------------------------------------------------------------------------------------
# sketches (src/sketch.jl)

def _op(A, trans):
    return A if trans == "n" else A.conj().T


def _orthrows(B):
    """orthrows!(thin=false): LQ + explicit Q; rows beyond min(m,n) zeroed (src/util.jl:88-99)."""
    m, n = B.shape
    k = min(m, n)
    q, _ = np.linalg.qr(B.T, mode="reduced")  # B' = QR  <=>  B = R'Q'
    out = np.zeros_like(B)
    out[:k, :] = q.T[:k, :]
    return out


def _orthcols(B):
    m, n = B.shape
    k = min(m, n)
    q, _ = np.linalg.qr(B, mode="reduced")
    out = np.zeros_like(B)
    out[:, :k] = q[:, :k]
    return out


def ishermitian(A) -> bool:
    A = np.asarray(A)
    return A.shape[0] == A.shape[1] and np.array_equal(A, A.conj().T)


def sketch(side: str, trans: str, A: np.ndarray, order: int, opts: LRAOptions, rand: Rand,
           rnd: int = 0) -> np.ndarray:
    """sketch (src/sketch.jl:35-50) for the four sketch types, left/right, :n/:c."""
    if opts.sketch == "randn":
        return sketch_randn(side, trans, A, order, opts, rand, rnd)
    m, n = A.shape
    if side == "left":
        X = _op(A, trans)
        mm = X.shape[0]
        if opts.sketch == "srft":
            d, idx = rand.srft(rnd, mm, order)
            return srft_left(X, order, d, idx)
        if opts.sketch == "sub":
            return subset_left(X, order, rand.subset(rnd, mm, order))
        if opts.sketch == "sprn":
            r, s = rand.sprn(rnd, mm, order)
            return sprn_left(X, order, r, s)
    else:
        # right sketches: A*S = (S'*A')' with the same structure along the other dimension
        X = _op(A, trans).conj().T
        return sketch("left", "n", X, order, opts, rand, rnd).conj().T
    raise ValueError(opts.sketch)


def sketch_randn(side, trans, A, order, opts, rand, rnd=0):
    """sketch_randn_{ln,lc,rn,rc} (src/sketch.jl:129-220) with injected Gaussian."""
    m, n = A.shape
    isherm = ishermitian(A)
    if side == "left":
        X = _op(A, trans)  # B = S * X
        S = rand.gaussian(rnd, order, X.shape[0])
        Bp = S @ X
        for _ in range(opts.sketch_randn_niter):
            Bp = _orthrows(Bp)
            Bq = Bp @ X.conj().T
            if isherm:
                Bp, Bq = Bq, Bp
            else:
                Bq = _orthrows(Bq)
                Bp = Bq @ X
        return np.asfortranarray(Bp)
    X = _op(A, trans)  # B = X * S
    S = rand.gaussian(rnd, X.shape[1], order)
    Bp = X @ S
    for _ in range(opts.sketch_randn_niter):
        Bp = _orthcols(Bp)
        Bq = X.conj().T @ Bp
        if isherm:
            Bp, Bq = Bq, Bp
        else:
            Bq = _orthcols(Bq)
            Bp = X @ Bq
    return np.asfortranarray(Bp)


@dataclasses.dataclass
class PQRFactors:
    Q: Optional[np.ndarray]
    R: Optional[np.ndarray]
    p: np.ndarray  # 1-based
    k: int
    T: Optional[np.ndarray]
    rounds: int = 1
    gap: Optional[np.ndarray] = None  # relative pivot gaps of the accepted QRCP


def pqrback_postproc(B, p, tau, k, opts: LRAOptions, use_openblas: bool = False) -> PQRFactors:
    """pqrback_postproc (src/pqr.jl:420-436); use_openblas solves T with OpenBLAS dtrsm."""
    retq = "q" in opts.pqrfact_retval
    retr = "r" in opts.pqrfact_retval
    rett = "t" in opts.pqrfact_retval
    n = B.shape[1]
    maxdet = 0 < k < n and opts.maxdet_tol >= 0
    Q = None
    if retq:
        Q = _orgqr(B[:, :k], tau[:k]) if k > 0 else np.zeros((B.shape[0], 0))
    R = np.triu(B[:k, :]) if (retr or rett or maxdet) else None
    T = (maxdet_t_blas(R) if use_openblas else maxdet_t(R)) if (rett or maxdet) else None
    p = np.array(p, dtype=np.int64)
    if maxdet:
        p, T, R1, it = maxdet_swapcols(p, T, opts.maxdet_tol, opts.maxdet_niter,
                                       R[:, :k] if retr else None)
        if it > 0:
            # qrfact!(R1) and Q*Qr (gemqrt) / R2 = R1*T (src/pqr.jl:467-475)
            if retr:
                qr_, r_ = np.linalg.qr(R1)
                if retq:
                    Q = Q @ qr_
                R = np.zeros_like(R)
                R[:, :k] = np.triu(r_)
                R[:, k:] = np.triu(r_) @ T
            # without retr, R1 is left untouched by the swaps and is upper triangular, so
            # its Householder QR is the identity and gemqrt leaves Q unchanged.
        if not retr:
            R = None
    return PQRFactors(Q, R, p, k, T)


def _orgqr(V, tau):
    """Explicit Q from Householder vectors below the diagonal of V (LAPACK orgqr)."""
    m, k = V.shape
    Q = np.eye(m, k)
    for j in range(k - 1, -1, -1):
        v = np.zeros(m)
        v[j] = 1.0
        v[j + 1:] = V[j + 1:, j]
        Q[j:, :] -= tau[j] * np.outer(v[j:], v[j:] @ Q[j:, :])
    return Q


def pqrfact_backend(A, opts: LRAOptions) -> PQRFactors:
    """pqrfact_backend! (src/pqr.jl:343-346)."""
    B, p, tau, k = geqp3_adap(A, opts.rank, opts.atol, opts.rtol, opts.nb)
    return pqrback_postproc(B, p, tau, k, opts)


def sketchfact(side: str, trans: str, A, opts: LRAOptions, rand: Rand,
               use_openblas: bool = False) -> PQRFactors:
    """sketchfact + sketchfact_{randn,srft,sub,sprn} adaptive drivers (src/sketch.jl:52-66,
    223-240, 313-330, 545-562, 674-690)."""
    sampfn = {"randn": opts.sketchfact_randn_samp, "srft": opts.sketchfact_srft_samp,
              "sub": opts.sketchfact_sub_samp, "sprn": (1, 0)}[opts.sketch]
    if opts.sketchfact_adap or opts.rank < 0:
        n = opts.nb
        opts_ = opts.copy(maxdet_tol=-1.0)
        rnd = 0
        while True:
            order = samp(sampfn, n)
            B = sketch(side, trans, A, order, opts, rand, rnd)
            if use_openblas:
                Bf, p, tau, k = geqp3_adap(B, opts_.rank, opts_.atol, opts_.rtol, opts_.nb,
                                           use_openblas=True)
                gap = None
            else:
                Bf, p, tau, k, gap = geqp3_adap(B, opts_.rank, opts_.atol, opts_.rtol,
                                                opts_.nb, want_gap=True)
            if k < n:
                F = pqrback_postproc(Bf, p, tau, k, opts, use_openblas)
                F.rounds = rnd + 1
                F.gap = gap
                return F
            n *= 2
            rnd += 1
    order = samp(sampfn, opts.rank)
    B = sketch(side, trans, A, order, opts, rand, 0)
    Bf, p, tau, k = geqp3_adap(B, opts.rank, opts.atol, opts.rtol, opts.nb,
                               use_openblas=use_openblas)
    return pqrback_postproc(Bf, p, tau, k, opts, use_openblas)


# ----------------------------------------------------------------------------------------
# factorizations

@dataclasses.dataclass
class IDPackedV:
    sk: np.ndarray  # 1-based
    rd: np.ndarray  # 1-based
    T: np.ndarray
    rounds: int = 1
    gap: Optional[np.ndarray] = None

    @property
    def k(self):
        return len(self.sk)

    def matrix(self) -> np.ndarray:
        """_Matrix(:n, V) = [I T] P' (src/id.jl:32-43): V[:,sk] = I, V[:,rd] = T."""
        k = len(self.sk)
        n = k + len(self.rd)
        M = np.zeros((k, n), dtype=self.T.dtype if self.T.size else np.float64)
        M[np.arange(k), self.sk - 1] = 1.0
        if len(self.rd):
            M[:, self.rd - 1] = self.T
        return M


def idfact(trans: str, A, opts: LRAOptions = None, rand: Rand = None,
           use_openblas: bool = False, **kw) -> IDPackedV:
    """idfact (src/id.jl:434-447)."""
    opts = (opts or LRAOptions()).copy(**kw)
    opts.pqrfact_retval = "t"
    chkopts(opts)
    rand = rand or Rand(0)
    if opts.sketch == "none":
        F = pqrfact_backend(_op(np.asarray(A, dtype=np.float64), trans).copy(order="F"), opts)
    else:
        F = sketchfact("left", trans, np.asarray(A, dtype=np.float64), opts, rand,
                       use_openblas=use_openblas)
    k = F.k
    T = F.T if F.T is not None else np.zeros((k, 0))
    return IDPackedV(F.p[:k].copy(), F.p[k:].copy(), T, F.rounds, F.gap)


def getcols(trans, A, cols):
    """getcols (src/util.jl:25-29); cols 1-based."""
    if trans == "n":
        return np.asfortranarray(A[:, cols - 1])
    return np.asfortranarray(A[cols - 1, :].conj().T)


@dataclasses.dataclass
class PartialQR:
    Q: np.ndarray
    R: np.ndarray
    p: np.ndarray

    def matrix(self):
        M = self.Q @ self.R
        out = np.empty_like(M)
        out[:, self.p - 1] = M
        return out


def pqrr(R, T):
    """pqrr (src/pqr.jl:330-340): [R  R*T]."""
    k = R.shape[0]
    return np.hstack([R, np.triu(R[:, :k]) @ T])


def prange(trans: str, A, opts: LRAOptions = None, rand: Rand = None, rand2: Rand = None,
           **kw) -> np.ndarray:
    """prange (src/prange.jl:14-77) and prange_sub (:65-78).  trans "n"/"c": randn takes the
    right-side sketchfact's Q (pqrfact_retval "q"), sub the QR of the left sketchfact's pivot
    columns, none pqrfact(op(A))[:Q]; "b" (square A): both right sketchfacts (`rand` for A',
    `rand2` for A), B = [Q_r R_r, Q_c R_c], then pqrfact_backend!(B)'s Q."""
    opts = (opts or LRAOptions()).copy(**kw)
    chkopts(opts)
    rand = rand or Rand(0)
    A = np.asarray(A, dtype=np.float64)
    if trans == "b":
        if A.shape[0] != A.shape[1]:
            raise ValueError("prange(:b) needs a square matrix")
        if ishermitian(A):
            return prange("n", A, opts, rand)
        o = opts.copy(pqrfact_retval="qr")
        if opts.sketch == "none":
            Fr = pqrfact_backend(np.asfortranarray(A.T), o)
            Fc = pqrfact_backend(np.asfortranarray(A), o)
        else:
            Fr = sketchfact("right", "c", A, o, rand)
            Fc = sketchfact("right", "n", A, o, rand2 or Rand(1))
        kr, kc = Fr.k, Fc.k
        B = np.hstack([Fr.Q @ np.triu(Fr.R[:kr, :kr]), Fc.Q @ np.triu(Fc.R[:kc, :kc])])
        return pqrfact_backend(np.asfortranarray(B), opts.copy(pqrfact_retval="q")).Q
    o = opts.copy(pqrfact_retval="q")
    if opts.sketch == "none":
        return pqrfact_backend(np.asfortranarray(_op(A, trans)), o).Q
    if opts.sketch == "sub":
        F = sketchfact("left", trans, A, opts.copy(pqrfact_retval=""), rand)
        cols = np.asarray(F.p[:F.k], dtype=np.int64) - 1
        B = _op(A, trans)[:, cols]
        return np.linalg.qr(B)[0] if F.k > 0 else np.zeros((B.shape[0], 0))
    return sketchfact("right", trans, A, o, rand).Q


def pqrfact(trans: str, A, opts: LRAOptions = None, rand: Rand = None, **kw):
    """pqrfact (src/pqr.jl:290-307)."""
    opts = (opts or LRAOptions()).copy(**kw)
    chkopts(opts)
    A = np.asarray(A, dtype=np.float64)
    if opts.sketch == "none":
        F = pqrfact_backend(_op(A, trans).copy(order="F"), opts)
        if "q" in opts.pqrfact_retval and "r" in opts.pqrfact_retval and "t" not in opts.pqrfact_retval:
            return PartialQR(F.Q, F.R, F.p)
        return F
    V = idfact(trans, A, opts, rand)
    Q, R = np.linalg.qr(getcols(trans, A, V.sk))
    retq, retr, rett = ("q" in opts.pqrfact_retval, "r" in opts.pqrfact_retval,
                        "t" in opts.pqrfact_retval)
    p = np.concatenate([V.sk, V.rd])
    if retq and retr and not rett:
        return PartialQR(Q, pqrr(R, V.T), p)
    return PQRFactors(Q if retq else None, pqrr(R, V.T) if retr else None, p, V.k,
                      V.T if rett else None)


@dataclasses.dataclass
class PartialSVD:
    U: np.ndarray
    S: np.ndarray
    Vt: np.ndarray

    def matrix(self):
        return (self.U * self.S) @ self.Vt


def psvdrank(s, opts: LRAOptions) -> int:
    """psvdrank (src/psvd.jl:301-308)."""
    k = len(s)
    if k == 0:
        return 0
    ptol = max(opts.atol, opts.rtol * s[0])
    for i in range(1, k):
        if s[i] <= ptol:
            return i
    return k


def psvdfact(A, opts: LRAOptions = None, rand: Rand = None, **kw) -> PartialSVD:
    """psvdfact (src/psvd.jl:238-272)."""
    opts = opts or LRAOptions()
    opts = opts.copy(**kw) if kw else opts
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    if m >= n:
        V = idfact("n", A, opts, rand)
        Q, R = np.linalg.qr(getcols("n", A, V.sk))
        Ut, s, Vh = np.linalg.svd(R @ V.matrix(), full_matrices=False)
        k = psvdrank(s, opts)
        return PartialSVD(Q @ Ut[:, :k], s[:k], Vh[:k, :])
    V = idfact("c", A, opts, rand)
    Q, R = np.linalg.qr(getcols("c", A, V.sk))
    Ut, s, Vh = np.linalg.svd(V.matrix().conj().T @ R.conj().T, full_matrices=False)
    k = psvdrank(s, opts)
    return PartialSVD(Ut[:, :k], s[:k], Vh[:k, :] @ Q.conj().T)


def psvdvals(A, opts: LRAOptions = None, rand: Rand = None, **kw):
    """psvdvals (src/psvd.jl:274-290)."""
    opts = opts or LRAOptions()
    opts = opts.copy(**kw) if kw else opts
    A = np.asarray(A, dtype=np.float64)
    m, n = A.shape
    if m >= n:
        V = idfact("n", A, opts, rand)
        Q, R = np.linalg.qr(getcols("n", A, V.sk))
        s = np.linalg.svd(R @ V.matrix(), compute_uv=False)
    else:
        V = idfact("c", A, opts, rand)
        Q, R = np.linalg.qr(getcols("c", A, V.sk))
        s = np.linalg.svd(V.matrix().conj().T @ R.conj().T, compute_uv=False)
    k = psvdrank(s, opts)
    return s[:k]


@dataclasses.dataclass
class PartialHermEigen:
    values: np.ndarray
    vectors: np.ndarray

    def matrix(self):
        return (self.vectors * self.values) @ self.vectors.conj().T


def hermitianize(B):
    """hermitianize!(B, :U) (src/util.jl:36-49): average into the upper triangle."""
    B = B.copy()
    n = B.shape[0]
    for j in range(n):
        for i in range(j + 1):
            B[i, j] = 0.5 * (B[i, j] + np.conj(B[j, i]))
    return np.triu(B) + np.triu(B, 1).conj().T


def pheigrank1(w, opts: LRAOptions, wmax: float) -> int:
    k = len(w)
    k = min(opts.rank, k) if opts.rank >= 0 else k
    ptol = max(opts.atol, opts.rtol * wmax)
    for i in range(1, k):
        if abs(w[i]) <= ptol:
            return i
    return k


def pheigrank(w, opts: LRAOptions):
    """pheigrank (src/pheig.jl:322-339)."""
    n = len(w)
    wmax = max(abs(w[0]), abs(w[n - 1]))
    lo = int(np.searchsorted(w, 0.0, "left"))
    hi = int(np.searchsorted(w, 0.0, "right"))
    kn = pheigrank1(w[:lo], opts, wmax)
    kp = pheigrank1(w[hi:][::-1], opts, wmax)
    return kn, kp


def pheigorth(values, vectors, opts: LRAOptions):
    """pheigorth! (src/pheig.jl:342-364): Gram-Schmidt inside eigenvalue clusters."""
    n = len(values)
    a = 0
    while a < n:
        va = values[a]
        b = a + 1
        while b < n:
            vb = values[b]
            if 2 * abs((va - vb) / (va + vb)) > opts.pheig_orthtol:
                break
            b += 1
        b -= 1
        for i in range(a, b + 1):
            for j in range(i + 1, b + 1):
                vectors[:, j] -= np.dot(vectors[:, i].conj(), vectors[:, j]) * vectors[:, i]
        a = b + 1
    return vectors


def pheigfact(A, opts: LRAOptions = None, rand: Rand = None, **kw) -> PartialHermEigen:
    """pheigfact (src/pheig.jl:276-296)."""
    opts = opts or LRAOptions()
    opts = opts.copy(**kw) if kw else opts
    A = np.asarray(A, dtype=np.float64)
    V = idfact("n", A, opts, rand)
    Q, R = np.linalg.qr(V.matrix().conj().T)
    sk = V.sk - 1
    B = R @ (A[np.ix_(sk, sk)] @ R.conj().T)
    w, X = np.linalg.eigh(hermitianize(B), UPLO="U")
    kn, kp = pheigrank(w, opts)
    n = len(w)
    if kn + kp < n:
        idx = np.concatenate([np.arange(kn), np.arange(n - kp, n)])
        w, X = w[idx], X[:, idx]
    X = pheigorth(w, X.copy(), opts)
    return PartialHermEigen(w, Q @ X)


def pheigvals(A, opts: LRAOptions = None, rand: Rand = None, **kw):
    """pheigvals (src/pheig.jl:298-311)."""
    opts = opts or LRAOptions()
    opts = opts.copy(**kw) if kw else opts
    A = np.asarray(A, dtype=np.float64)
    V = idfact("n", A, opts, rand)
    Q, R = np.linalg.qr(V.matrix().conj().T)
    sk = V.sk - 1
    B = R @ (A[np.ix_(sk, sk)] @ R.conj().T)
    w = np.linalg.eigvalsh(hermitianize(B), UPLO="U")
    kn, kp = pheigrank(w, opts)
    n = len(w)
    if kn + kp < n:
        return np.concatenate([w[:kn], w[n - kp:]])
    return w


def curfact(A, opts: LRAOptions = None, rand: Rand = None, rand2: Rand = None, **kw):
    """curfact (src/cur.jl:532-564); returns (rows, cols) 1-based (Hermitian/symmetric:
    rows == cols).  rand2, when given, supplies the second sketchfact's randomness (default:
    rand, continuing as the reference's global generator would)."""
    opts = (opts or LRAOptions()).copy(**kw)
    opts.pqrfact_retval = ""
    chkopts(opts)
    A = np.asarray(A, dtype=np.float64)
    rand = rand or Rand(0)
    if opts.sketch == "none":
        return curfact_none(A, opts)
    if ishermitian(A):
        F = sketchfact("left", "n", A, opts, rand)
        cols = F.p[:F.k]
        return cols, cols
    m, n = A.shape
    rand2 = rand2 or rand
    if m >= n:
        F = sketchfact("left", "c", A, opts, rand)
        rows = F.p[:F.k]
        F = sketchfact("left", "n", A[rows - 1, :], opts, rand2)
        cols = F.p[:F.k]
    else:
        F = sketchfact("left", "n", A, opts, rand)
        cols = F.p[:F.k]
        F = sketchfact("left", "c", A[:, cols - 1], opts, rand2)
        rows = F.p[:F.k]
    k = min(len(rows), len(cols))
    return rows[:k], cols[:k]


def cur_build(A, rows, cols):
    """CUR(A, CURPackedU) (src/cur.jl:85-93): C = A[:, cols], R = A[rows, :], and U =
    PartialSVD(V, 1 ./ sigma, U') from the SVD of W = A[rows, cols] (1-based skeletons).
    Returns (C, (V, sinv, Ut), R)."""
    A = np.asarray(A, dtype=np.float64)
    r = np.asarray(rows, dtype=np.int64) - 1
    c = np.asarray(cols, dtype=np.int64) - 1
    C = A[:, c]
    R = A[r, :]
    Uw, s, Vt = np.linalg.svd(A[np.ix_(r, c)])
    with np.errstate(divide="ignore"):
        sinv = 1.0 / s
    return C, (Vt.T, sinv, Uw.T), R


def curfact_none(A, opts: LRAOptions):
    """curfact_none (src/cur.jl:574-606)."""
    if ishermitian(A):
        F = pqrfact_backend(A.copy(order="F"), opts)
        cols = F.p[:F.k]
        return cols, cols
    m, n = A.shape
    if m >= n:
        F = pqrfact_backend(np.asfortranarray(A.conj().T), opts)
        rows = F.p[:F.k]
        F = pqrfact_backend(np.asfortranarray(A[rows - 1, :]), opts)
        cols = F.p[:F.k]
    else:
        F = pqrfact_backend(A.copy(order="F"), opts)
        cols = F.p[:F.k]
        F = pqrfact_backend(np.asfortranarray(A[:, cols - 1].conj().T), opts)
        rows = F.p[:F.k]
    k = min(len(rows), len(cols))
    return rows[:k], cols[:k]


def snorm(A: Callable | np.ndarray, n: int = None, opts: LRAOptions = None, x0=None,
          isherm: bool = None, AH: Callable = None, **kw) -> float:
    """snorm (src/snorm.jl:14-45): power iteration on A'A (A if Hermitian).  A is a dense
    matrix or a callable x -> A x (then AH must be x -> A' x and n the column count)."""
    opts = opts or LRAOptions()
    opts = opts.copy(**kw) if kw else opts
    if not callable(A):
        M = np.asarray(A, dtype=np.float64)
        isherm = ishermitian(M) if isherm is None else isherm
        n = M.shape[1]
        A, AH = (lambda x: M @ x), (lambda x: M.conj().T @ x)
    isherm = bool(isherm)
    xn = np.array(x0, dtype=np.float64) if x0 is not None else Rand(0).start(n)
    xnrm = np.linalg.norm(xn)
    s, t, niter = 1.0, 0.0, 0
    while s > 0 and abs(s - t) > max(opts.atol, t * opts.rtol):
        if niter == opts.snorm_niter:
            break
        niter += 1
        xn = xn * (1.0 / xnrm)
        if isherm:
            xn = A(xn)
        else:
            xn = AH(A(xn))
        xnrm = np.linalg.norm(xn)
        t = s
        s = xnrm if isherm else np.sqrt(xnrm)
    return float(s)


def snormdiff(A, B, opts: LRAOptions = None, x0=None, **kw) -> float:
    """snormdiff (src/snorm.jl:50-53) = snorm(A - B)."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    return snorm(A - B, opts=opts, x0=x0, **kw)


def id_residual(A, V: IDPackedV, trans: str = "n") -> np.ndarray:
    """A - A[:,sk] * [I T] P' (the ID's reconstruction error matrix)."""
    X = _op(np.asarray(A, dtype=np.float64), trans)
    return X - X[:, V.sk - 1] @ V.matrix()


def hilbert(m: int, n: int = None) -> np.ndarray:
    """Hilbert matrix 1/(i+j-1) (src/matrixlib.jl:39-48)."""
    n = m if n is None else n
    i = np.arange(1, m + 1)[:, None]
    j = np.arange(1, n + 1)[None, :]
    return np.asfortranarray(1.0 / (i + j - 1))


def fourier(x, y) -> np.ndarray:
    """matrixlib(:fourier, x, y) (src/matrixlib.jl:25-33): A[i,j] = exp(-2 pi i x_i y_j), the
    complex test matrix of test/id.jl:7,24."""
    return np.asfortranarray(np.exp(-2j * np.pi * np.outer(x, y)))


def fourier_real(x, y) -> np.ndarray:
    """Real part of matrixlib(:fourier, x, y) (src/matrixlib.jl:25-33), as the reference
    tests use for Float64 (test/id.jl:7,24)."""
    return np.asfortranarray(np.cos(-2 * np.pi * np.outer(x, y)))
