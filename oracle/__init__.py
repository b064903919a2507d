"""Test infrastructure, not product code: ctypes bindings to the FP64 oracle.

The oracle (``pc_oracle.c``) is a CPU restatement of gpucsl's PC-stable
skeleton discovery.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.  The
product (``gpucsl_amd``) never imports it.

``ref()`` returns bindings to ``_ref/libref.so``: the reference's own
host-compilable helpers, built from the reference checkout by ``Makefile``.
It is absent on machines without the checkout; tests that need it skip.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref.so"

_P = ct.c_void_p
_i64 = ct.c_int64


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference checkout exists)."""
    args = ["make", "-C", str(HERE)] + (["-B"] if force else []) + ["all"]
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ct.CDLL(str(LIB))
        L.or_binomial.restype = _i64
        L.or_binomial.argtypes = [_i64, _i64]
        L.or_unrank.argtypes = [ct.c_int, ct.c_int, _i64, _P]
        L.or_poz.restype = ct.c_double
        L.or_poz.argtypes = [ct.c_double]
        L.or_pchisq.restype = ct.c_double
        L.or_pchisq.argtypes = [ct.c_double, ct.c_int]
        L.or_gauss_z.restype = ct.c_double
        L.or_gauss_z.argtypes = [_P, ct.c_int, ct.c_int, ct.c_int, _P, ct.c_int]
        L.or_gauss_z0.restype = ct.c_double
        L.or_gauss_z0.argtypes = [ct.c_double]
        L.or_corrcoef.argtypes = [_P, _i64, ct.c_int, _P]
        L.or_compact.argtypes = [_P, ct.c_int, _P]
        L.or_gauss_skeleton.restype = ct.c_int
        L.or_gauss_skeleton.argtypes = [_P, ct.c_int, _P, ct.c_int, _P, _P, _P, _P, _P]
        L.or_gauss_skeleton_ex.restype = ct.c_int
        L.or_gauss_skeleton_ex.argtypes = [_P, ct.c_int, _P, ct.c_int, _i64, ct.c_int, _P, _P, _P, _P, _P]
        L.or_pmax_from_zmin.argtypes = [_P, ct.c_int, _i64, _P]
        L.or_chi2_table.restype = ct.c_double
        L.or_chi2_table.argtypes = [_P, _P, _P, _P, ct.c_int, ct.c_int, ct.c_int, ct.c_int]
        L.or_disc_chi2.restype = ct.c_double
        L.or_disc_chi2.argtypes = [_P, _i64, _P, ct.c_int, ct.c_int, _P, ct.c_int, _P, _P]
        L.or_disc_skeleton.restype = ct.c_int
        L.or_disc_skeleton.argtypes = [_P, _i64, ct.c_int, _P, ct.c_double, ct.c_int, _P, _P, _P,
                                       _P, _P]
        L.or_num_threads.restype = ct.c_int
        L.or_set_threads.restype = None
        L.or_set_threads.argtypes = [ct.c_int]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def binomial(n: int, k: int) -> int:
    return int(lib().or_binomial(n, k))


def unrank(N: int, l: int, rank: int) -> np.ndarray:
    out = np.zeros(max(l, 1), np.int32)
    lib().or_unrank(N, l, rank, _ptr(out))
    return out[:l]


def poz(z: float) -> float:
    return float(lib().or_poz(z))


def pchisq(x: float, df: int) -> float:
    return float(lib().or_pchisq(x, df))


def gauss_z(C: np.ndarray, vi: int, vj: int, S) -> float:
    C = np.ascontiguousarray(C, np.float64)
    S = np.ascontiguousarray(S, np.int32)
    return float(lib().or_gauss_z(_ptr(C), C.shape[1], vi, vj, _ptr(S), len(S)))


def gauss_z0(r: float) -> float:
    return float(lib().or_gauss_z0(r))


def corrcoef(X: np.ndarray) -> np.ndarray:
    X = np.ascontiguousarray(X, np.float64)
    n, p = X.shape
    C = np.empty((p, p), np.float64)
    lib().or_corrcoef(_ptr(X), n, p, _ptr(C))
    return C


def compact(G: np.ndarray) -> np.ndarray:
    G = np.ascontiguousarray(G, np.uint16)
    p = G.shape[0]
    Gc = np.zeros((p, p), np.int32)
    lib().or_compact(_ptr(G), p, _ptr(Gc))
    return Gc


def thresholds(n: int, alpha: float, levels: int = 50) -> np.ndarray:
    """|Phi^-1(alpha/2)| / sqrt(n - l - 4) for l < min(50, n-3) - 2, else 0
    (gpucsl/pc/helpers.py:42-50, including its off-by-one)."""
    from scipy.stats import norm

    thr = np.zeros(levels, np.float64)
    q = abs(norm.ppf(alpha / 2, loc=0, scale=1))
    for i in range(1, min(levels, n - 3) - 1):
        thr[i - 1] = abs(q / np.sqrt(n - i - 3))
    return thr


def pcalg_thresholds(n: int, alpha: float, levels: int = 50) -> np.ndarray:
    """pcalg's |Phi^-1(alpha/2)| / sqrt(n - l - 3) (0 where n - l - 3 <= 0)."""
    from scipy.stats import norm

    q = abs(norm.ppf(alpha / 2))
    return np.array([q / np.sqrt(n - l - 3) if n - l - 3 > 0 else 0.0 for l in range(levels)])


def gauss_skeleton(C: np.ndarray, n: int, alpha: float, max_level: int, thr=None, pcalg_compat: bool = False) -> dict:
    """FP64 Gaussian PC-stable skeleton on a correlation matrix.  pcalg_compat:
    pcalg's thresholds, pMax (largest p-value of any test) and one-way sepsets."""
    C = np.ascontiguousarray(C, np.float64)
    p = C.shape[0]
    if thr is None:
        thr = pcalg_thresholds(n, alpha) if pcalg_compat else thresholds(n, alpha)
    thr = np.ascontiguousarray(thr, np.float64)
    L = max(max_level, 1)
    G = np.zeros((p, p), np.uint8)
    sep = np.empty((p, p, L), np.int32)
    zmin = np.empty((p, p), np.float32)
    ex = np.zeros(max_level + 1, np.int64)
    tot = np.zeros(max_level + 1, np.int64)
    last = lib().or_gauss_skeleton_ex(_ptr(C), p, _ptr(thr), max_level, ct.c_int64(n), 1 if pcalg_compat else 0,
                                      _ptr(G), _ptr(sep), _ptr(zmin), _ptr(ex), _ptr(tot))
    pmax = np.empty((p, p), np.float32)
    # pcalg mode: zmin already carries sqrt(n - |S| - 3); n = 4 makes the scale 1
    lib().or_pmax_from_zmin(_ptr(zmin), p, 4 if pcalg_compat else n, _ptr(pmax))
    return dict(skeleton=G.astype(np.uint16), sepsets=sep[:, :, :max_level], zmin=zmin,
                pmax=pmax, executed=ex, total=tot, last_level=last)


def pmax_from_zmin(zmin: np.ndarray, n: int) -> np.ndarray:
    zmin = np.ascontiguousarray(zmin, np.float32)
    p = zmin.shape[0]
    out = np.empty((p, p), np.float32)
    lib().or_pmax_from_zmin(_ptr(zmin), p, n, _ptr(out))
    return out


def chi2_table(N, Nis, Njs, Ns, dim_s, di, dj, sequential=False) -> float:
    arrs = [np.ascontiguousarray(a, np.uint32) for a in (N, Nis, Njs, Ns)]
    return float(lib().or_chi2_table(*[_ptr(a) for a in arrs], dim_s, di, dj, int(sequential)))


def disc_chi2(data_colmajor: np.ndarray, dims: np.ndarray, vi: int, vj: int, S):
    """(chi2, df, table) for one test on column-major (p, n) uint8 data."""
    d = np.ascontiguousarray(data_colmajor, np.uint8)
    dims = np.ascontiguousarray(dims, np.uint8)
    S = np.ascontiguousarray(S, np.int32)
    maxd = int(dims.max())
    work = np.zeros(maxd ** (len(S) + 2) * 4 + 16, np.uint32)
    df = ct.c_int(0)
    chi2 = lib().or_disc_chi2(_ptr(d), d.shape[1], _ptr(dims), vi, vj, _ptr(S), len(S),
                              _ptr(work), ct.byref(df))
    return float(chi2), int(df.value), work


def disc_skeleton(data_colmajor: np.ndarray, dims: np.ndarray, alpha: float, max_level: int) -> dict:
    d = np.ascontiguousarray(data_colmajor, np.uint8)
    dims = np.ascontiguousarray(dims, np.uint8)
    p, n = d.shape
    L = max(max_level, 1)
    G = np.zeros((p, p), np.uint8)
    sep = np.empty((p, p, L), np.int32)
    pmax = np.empty((p, p), np.float32)
    ex = np.zeros(max_level + 1, np.int64)
    tot = np.zeros(max_level + 1, np.int64)
    last = lib().or_disc_skeleton(_ptr(d), n, p, _ptr(dims), alpha, max_level, _ptr(G), _ptr(sep),
                                  _ptr(pmax), _ptr(ex), _ptr(tot))
    pm = pmax.copy()
    pm[np.tril_indices(p)] = -1.0
    return dict(skeleton=G.astype(np.uint16), sepsets=sep[:, :, :max_level], pmax=pm,
                executed=ex, total=tot, last_level=last)


def num_threads() -> int:
    return int(lib().or_num_threads())


def set_threads(k: int) -> None:
    """Threads of the following oracle calls (for timing); 0 restores the default."""
    lib().or_set_threads(int(k))


# --------------------------------------------------------------------------
# reference helpers compiled for the host (oracle/_ref)
# --------------------------------------------------------------------------
_ref = None


def ref():
    """Bindings to oracle/_ref/libref.so, or None when it was not built."""
    global _ref
    if _ref is None:
        if not REF_LIB.exists():
            return None
        R = ct.CDLL(str(REF_LIB))
        R.ref_binomial.restype = ct.c_int
        R.ref_binomial.argtypes = [ct.c_int, ct.c_int]
        R.ref_poz.restype = ct.c_double
        R.ref_poz.argtypes = [ct.c_double]
        R.ref_pchisq.restype = ct.c_double
        R.ref_pchisq.argtypes = [ct.c_double, ct.c_int]
        R.ref_sepset_indices.restype = ct.c_int
        R.ref_sepset_indices.argtypes = [ct.c_int, ct.c_int, ct.c_int, _P]
        R.ref_gauss_z_packed.restype = ct.c_double
        R.ref_gauss_z_packed.argtypes = [ct.c_int, _P]
        for f in ("ref_inverse3", "ref_cholesky3", "ref_m2_inv3"):
            getattr(R, f).argtypes = [_P, _P]
        _ref = R
    return _ref
