"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

Test infrastructure only: the tests, the smoke check and the CPU-baseline leg
of the benchmark use it as the checker; the product never imports it.

Arrays are numpy uint64 with the reference's in-memory layout: an Fr/Fq
element is 4 words (Montgomery residue, little-endian), a G1 point 12 words
(Jacobian X, Y, Z), a G2 point 24 words (X.c0, X.c1, Y.c0, Y.c1, Z.c0, Z.c1).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

Q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
MONT = 1 << 256


def build() -> None:
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class Csr(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_size_t), ("cols", ctypes.c_size_t), ("nnz", ctypes.c_size_t),
                ("row_ptr", ctypes.c_void_p), ("col_idx", ctypes.c_void_p), ("values", ctypes.c_void_p)]


class Circuit(ctypes.Structure):
    _fields_ = [("public_inputs", ctypes.c_size_t), ("a", Csr), ("b", Csr), ("c", Csr)]


class Pk(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "a_query", "b_g1_query", "b_g2_query", "l_query", "h_query",
        "alpha_g1", "beta_g1", "beta_g2", "delta_g1", "delta_g2")]


def _declare(L: ctypes.CDLL) -> None:
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    for name in ("or_fr_mul", "or_fr_add", "or_fr_sub", "or_fq_mul", "or_fq_add", "or_fq_sub", "or_fq2_mul"):
        getattr(L, name).argtypes = [vp, vp, vp, sz]
    for name in ("or_fr_inv", "or_fr_to_mont", "or_fr_from_mont", "or_fq_to_mont", "or_fq_from_mont",
                 "or_fq2_sqr", "or_g1_to_affine", "or_g2_to_affine"):
        getattr(L, name).argtypes = [vp, vp, sz]
    L.or_mt19937_64.argtypes = [ctypes.c_uint64, vp, sz]
    L.or_fr_random.argtypes = [ctypes.c_uint64, vp, sz]
    L.or_g1_generator.argtypes = [vp]
    L.or_g2_generator.argtypes = [vp]
    L.or_g1_add.argtypes = [vp, vp, vp]
    L.or_g2_add.argtypes = [vp, vp, vp]
    L.or_g1_dbl.argtypes = [vp, vp]
    L.or_g1_neg.argtypes = [vp, vp]
    L.or_g1_mul.argtypes = [vp, vp, vp, sz]
    L.or_g2_mul.argtypes = [vp, vp, vp, sz]
    L.or_g1_on_curve.argtypes = [vp]
    L.or_g2_on_curve.argtypes = [vp]
    L.or_spmv.argtypes = [ctypes.POINTER(Csr), vp, vp]
    L.or_fft.argtypes = [vp, sz, ctypes.c_int]
    L.or_fft.restype = ctypes.c_int
    L.or_quotient.argtypes = [vp, vp, vp, sz, vp]
    L.or_quotient.restype = ctypes.c_int
    L.or_msm_g1.argtypes = [vp, vp, sz, vp]
    L.or_msm_g2.argtypes = [vp, vp, sz, vp]
    L.or_prove.argtypes = [ctypes.POINTER(Circuit), ctypes.POINTER(Pk), vp, vp, vp, vp, vp, vp]
    L.or_prove.restype = ctypes.c_int
    L.or_blind.argtypes = [ctypes.POINTER(Pk), vp, vp, vp, vp]
    L.or_msm_g1_mt.argtypes = [vp, vp, sz, vp, sz]
    L.or_msm_g2_mt.argtypes = [vp, vp, sz, vp, sz]
    L.or_prove_mt.argtypes = [ctypes.POINTER(Circuit), ctypes.POINTER(Pk), vp, vp, vp, vp, vp, vp, sz]
    L.or_prove_mt.restype = ctypes.c_int


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


# ---------------------------------------------------------------- conversions
def int_to_words(x: int, words: int = 4) -> list[int]:
    return [(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(words)]


def words_to_int(w) -> int:
    return sum(int(v) << (64 * i) for i, v in enumerate(w))


def to_mont_int(x: int, p: int) -> int:
    return x * MONT % p


def from_mont_int(x: int, p: int) -> int:
    return x * pow(MONT, -1, p) % p


def fr_array(values) -> np.ndarray:
    """Integers (canonical) -> Montgomery Fr array (n, 4)."""
    return np.array([int_to_words(to_mont_int(v % R, R)) for v in values], dtype=np.uint64).reshape(-1, 4)


_RINV_R = pow(MONT, -1, R)


def fr_ints(arr: np.ndarray) -> list[int]:
    a = np.ascontiguousarray(arr, dtype=np.uint64).reshape(-1, 4)
    w = [a[:, i].tolist() for i in range(4)]
    return [((w3 << 192) | (w2 << 128) | (w1 << 64) | w0) * _RINV_R % R
            for w0, w1, w2, w3 in zip(*w)]


# ------------------------------------------------------------------ wrappers
def mt19937_64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().or_mt19937_64(seed, _p(out), n)
    return out


def fr_random(seed: int, n: int) -> np.ndarray:
    out = np.zeros((n, 4), dtype=np.uint64)
    lib().or_fr_random(seed, _p(out), n)
    return out


def _bin(name: str, a: np.ndarray, b: np.ndarray, width: int = 4) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, width)
    b = np.ascontiguousarray(b, dtype=np.uint64).reshape(-1, width)
    out = np.zeros_like(a)
    getattr(lib(), name)(_p(a), _p(b), _p(out), a.shape[0])
    return out


def _un(name: str, a: np.ndarray, width: int = 4) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64).reshape(-1, width)
    out = np.zeros_like(a)
    getattr(lib(), name)(_p(a), _p(out), a.shape[0])
    return out


def fr_mul(a, b): return _bin("or_fr_mul", a, b)
def fr_add(a, b): return _bin("or_fr_add", a, b)
def fr_sub(a, b): return _bin("or_fr_sub", a, b)
def fq_mul(a, b): return _bin("or_fq_mul", a, b)
def fq_add(a, b): return _bin("or_fq_add", a, b)
def fq_sub(a, b): return _bin("or_fq_sub", a, b)
def fq2_mul(a, b): return _bin("or_fq2_mul", a, b, 8)
def fq2_sqr(a): return _un("or_fq2_sqr", a, 8)
def fr_inv(a): return _un("or_fr_inv", a)
def fr_to_mont(a): return _un("or_fr_to_mont", a)
def fr_from_mont(a): return _un("or_fr_from_mont", a)
def fq_to_mont(a): return _un("or_fq_to_mont", a)
def fq_from_mont(a): return _un("or_fq_from_mont", a)
def g1_to_affine(a): return _un("or_g1_to_affine", a, 12)
def g2_to_affine(a): return _un("or_g2_to_affine", a, 24)


def g1_generator() -> np.ndarray:
    out = np.zeros(12, dtype=np.uint64)
    lib().or_g1_generator(_p(out))
    return out


def g2_generator() -> np.ndarray:
    out = np.zeros(24, dtype=np.uint64)
    lib().or_g2_generator(_p(out))
    return out


def g1_add(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros(12, dtype=np.uint64)
    lib().or_g1_add(_p(a), _p(b), _p(out))
    return out


def g2_add(a, b):
    a = np.ascontiguousarray(a, dtype=np.uint64); b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.zeros(24, dtype=np.uint64)
    lib().or_g2_add(_p(a), _p(b), _p(out))
    return out


def g1_neg(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.zeros(12, dtype=np.uint64)
    lib().or_g1_neg(_p(a), _p(out))
    return out


def g1_mul(bases: np.ndarray, scalars: np.ndarray) -> np.ndarray:
    bases = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 12)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    if bases.shape[0] == 1 and scalars.shape[0] > 1:
        bases = np.ascontiguousarray(np.repeat(bases, scalars.shape[0], axis=0))
    out = np.zeros_like(bases)
    lib().or_g1_mul(_p(bases), _p(scalars), _p(out), bases.shape[0])
    return out


def g2_mul(bases: np.ndarray, scalars: np.ndarray) -> np.ndarray:
    bases = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 24)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    if bases.shape[0] == 1 and scalars.shape[0] > 1:
        bases = np.ascontiguousarray(np.repeat(bases, scalars.shape[0], axis=0))
    out = np.zeros_like(bases)
    lib().or_g2_mul(_p(bases), _p(scalars), _p(out), bases.shape[0])
    return out


def g1_on_curve(p) -> bool:
    p = np.ascontiguousarray(p, dtype=np.uint64)
    return bool(lib().or_g1_on_curve(_p(p)))


def g2_on_curve(p) -> bool:
    p = np.ascontiguousarray(p, dtype=np.uint64)
    return bool(lib().or_g2_on_curve(_p(p)))


def make_csr(rows: int, cols: int, row_ptr: np.ndarray, col_idx: np.ndarray, values: np.ndarray):
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.uint32)
    values = np.ascontiguousarray(values, dtype=np.uint64).reshape(-1, 4)
    c = Csr(rows, cols, values.shape[0], row_ptr.ctypes.data, col_idx.ctypes.data, values.ctypes.data)
    c._keep = (row_ptr, col_idx, values)
    return c


def spmv(rows, cols, row_ptr, col_idx, values, z) -> np.ndarray:
    c = make_csr(rows, cols, row_ptr, col_idx, values)
    z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros((rows, 4), dtype=np.uint64)
    lib().or_spmv(ctypes.byref(c), _p(z), _p(out))
    return out


FFT, IFFT, COSET_FFT, ICOSET_FFT = 0, 1, 2, 3


def fft(a: np.ndarray, kind: int) -> np.ndarray:
    out = np.ascontiguousarray(np.array(a, dtype=np.uint64).reshape(-1, 4))
    rc = lib().or_fft(_p(out), out.shape[0], kind)
    if rc != 0:
        raise ValueError("or_fft: unsupported size")
    return out


def quotient(az, bz, cz) -> np.ndarray:
    az = np.ascontiguousarray(az, dtype=np.uint64).reshape(-1, 4)
    bz = np.ascontiguousarray(bz, dtype=np.uint64).reshape(-1, 4)
    cz = np.ascontiguousarray(cz, dtype=np.uint64).reshape(-1, 4)
    h = np.zeros_like(az)
    if lib().or_quotient(_p(az), _p(bz), _p(cz), az.shape[0], _p(h)) != 0:
        raise ValueError("or_quotient: unsupported size")
    return h


def msm_g1(bases: np.ndarray, scalars: np.ndarray) -> np.ndarray:
    bases = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 12)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros(12, dtype=np.uint64)
    lib().or_msm_g1(_p(bases), _p(scalars), bases.shape[0], _p(out))
    return out


def msm_g2(bases: np.ndarray, scalars: np.ndarray) -> np.ndarray:
    bases = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 24)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros(24, dtype=np.uint64)
    lib().or_msm_g2(_p(bases), _p(scalars), bases.shape[0], _p(out))
    return out


def host_threads() -> int:
    """Host threads for the parallel oracle: the CPUs this process may run on,
    at most 16 (the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def msm_g1_mt(bases: np.ndarray, scalars: np.ndarray, threads: int | None = None) -> np.ndarray:
    """variable_base_msm: chunked BDLO12 over `threads` host threads."""
    bases = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 12)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros(12, dtype=np.uint64)
    lib().or_msm_g1_mt(_p(bases), _p(scalars), bases.shape[0], _p(out), threads or host_threads())
    return out


def msm_g2_mt(bases: np.ndarray, scalars: np.ndarray, threads: int | None = None) -> np.ndarray:
    bases = np.ascontiguousarray(bases, dtype=np.uint64).reshape(-1, 24)
    scalars = np.ascontiguousarray(scalars, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros(24, dtype=np.uint64)
    lib().or_msm_g2_mt(_p(bases), _p(scalars), bases.shape[0], _p(out), threads or host_threads())
    return out


def _circuit(circ) -> Circuit:
    """circ: object with public_inputs, rows, cols, and a/b/c as (row_ptr, col_idx, values)."""
    mats = [make_csr(circ.rows, circ.cols, *m) for m in (circ.a, circ.b, circ.c)]
    c = Circuit(circ.public_inputs, mats[0], mats[1], mats[2])
    c._keep = mats
    return c


def _pk(pk) -> Pk:
    arrs = {k: np.ascontiguousarray(getattr(pk, k), dtype=np.uint64) for k in (
        "a_query", "b_g1_query", "b_g2_query", "l_query", "h_query",
        "alpha_g1", "beta_g1", "beta_g2", "delta_g1", "delta_g2")}
    p = Pk(**{k: v.ctypes.data for k, v in arrs.items()})
    p._keep = arrs
    return p


def prove(circ, pk, witness: np.ndarray, intermediates: bool = False):
    """CpuBackend::prove.  Returns raw proof words (72) [, (az, bz, cz, h)]."""
    c = _circuit(circ)
    k = _pk(pk)
    w = np.ascontiguousarray(witness, dtype=np.uint64).reshape(-1, 4)
    raw = np.zeros(72, dtype=np.uint64)
    n = circ.rows
    inter = [np.zeros((n, 4), dtype=np.uint64) for _ in range(4)] if intermediates else [None] * 4
    ptrs = [(_p(a) if a is not None else None) for a in inter]
    if lib().or_prove(ctypes.byref(c), ctypes.byref(k), _p(w), _p(raw), *ptrs) != 0:
        raise ValueError("or_prove failed")
    return (raw, tuple(inter)) if intermediates else raw


def prove_mt(circ, pk, witness: np.ndarray, intermediates: bool = False, threads: int | None = None):
    """or_prove on several host threads (chunked MSMs); same results."""
    c = _circuit(circ)
    k = _pk(pk)
    w = np.ascontiguousarray(witness, dtype=np.uint64).reshape(-1, 4)
    raw = np.zeros(72, dtype=np.uint64)
    n = circ.rows
    inter = [np.zeros((n, 4), dtype=np.uint64) for _ in range(4)] if intermediates else [None] * 4
    ptrs = [(_p(a) if a is not None else None) for a in inter]
    if lib().or_prove_mt(ctypes.byref(c), ctypes.byref(k), _p(w), _p(raw), *ptrs, threads or host_threads()) != 0:
        raise ValueError("or_prove_mt failed")
    return (raw, tuple(inter)) if intermediates else raw


def blind(pk, raw: np.ndarray, r: np.ndarray, s: np.ndarray) -> np.ndarray:
    k = _pk(pk)
    raw = np.ascontiguousarray(raw, dtype=np.uint64)
    r = np.ascontiguousarray(r, dtype=np.uint64).reshape(4)
    s = np.ascontiguousarray(s, dtype=np.uint64).reshape(4)
    out = np.zeros(48, dtype=np.uint64)
    lib().or_blind(ctypes.byref(k), _p(raw), _p(r), _p(s), _p(out))
    return out
