"""ctypes loader of oracle/build/liboracle.so (the C restatement).

TEST INFRASTRUCTURE ONLY -- see oracle/__init__.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        i64, p = ctypes.c_int64, ctypes.c_void_p
        for suf in ("f32", "f64"):
            getattr(_lib, f"kmo_half_norm_{suf}").argtypes = [i64, i64, p, p]
            getattr(_lib, f"kmo_assign_{suf}").argtypes = [i64, i64, i64, p, p, p, p,
                                                          ctypes.c_int]
            f = getattr(_lib, f"kmo_lloyd_{suf}")
            f.argtypes = [i64, i64, i64, p, p, p, i64, ctypes.c_double, p, p,
                          ctypes.c_int]
            f.restype = i64
        _lib.kmo_max_threads.restype = ctypes.c_int
    return _lib


def _suffix(dtype) -> str:
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def max_threads() -> int:
    return lib().kmo_max_threads()


def half_norm(C_t: np.ndarray) -> np.ndarray:
    C_t = np.ascontiguousarray(C_t)
    hn = np.empty(C_t.shape[1], dtype=C_t.dtype)
    getattr(lib(), f"kmo_half_norm_{_suffix(C_t.dtype)}")(
        C_t.shape[0], C_t.shape[1], _ptr(C_t), _ptr(hn))
    return hn


def assign(X_t: np.ndarray, C_t: np.ndarray, hn: np.ndarray | None = None,
           nthreads: int | None = None) -> np.ndarray:
    X_t = np.ascontiguousarray(X_t)
    C_t = np.ascontiguousarray(C_t, dtype=X_t.dtype)
    hn = half_norm(C_t) if hn is None else np.ascontiguousarray(hn, dtype=X_t.dtype)
    labels = np.empty(X_t.shape[1], dtype=np.int64)
    getattr(lib(), f"kmo_assign_{_suffix(X_t.dtype)}")(
        X_t.shape[1], X_t.shape[0], C_t.shape[1], _ptr(X_t), _ptr(C_t),
        _ptr(hn), _ptr(labels), nthreads or max_threads())
    return labels


def lloyd(X_t: np.ndarray, w: np.ndarray, init_t: np.ndarray, tol: float,
          max_iter: int, nthreads: int | None = None):
    """Returns (labels, centroids_t, n_iter, inertia)."""
    X_t = np.ascontiguousarray(X_t)
    w = np.ascontiguousarray(w, dtype=X_t.dtype)
    C = np.array(init_t, dtype=X_t.dtype, copy=True, order="C")
    labels = np.empty(X_t.shape[1], dtype=np.int64)
    inertia = ctypes.c_double(0.0)
    it = getattr(lib(), f"kmo_lloyd_{_suffix(X_t.dtype)}")(
        X_t.shape[1], X_t.shape[0], C.shape[1], _ptr(X_t), _ptr(w), _ptr(C),
        max_iter, tol, _ptr(labels), ctypes.byref(inertia),
        nthreads or max_threads())
    return labels, C, int(it), float(inertia.value)
