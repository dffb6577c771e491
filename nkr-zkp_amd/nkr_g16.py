"""ctypes binding of libnkr_g16.so, the gfx950 Groth16 backend (C ABI in
include/nkr_g16.h).  Used by the tests, the smoke check and the benchmark.

There is no fallback: if the shared library is missing, importing the
functions raises.  Arrays are numpy uint64 in the reference's in-memory
layouts (Fr/Fq 4 words, G1 Jacobian 12 words, G2 Jacobian 24 words).
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "lib", "libnkr_g16.so")
HEADER = os.path.join(ROOT, "include", "nkr_g16.h")


class NkrError(RuntimeError):
    pass


class Circuit(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_uint64), ("cols", ctypes.c_uint64), ("public_inputs", ctypes.c_uint64),
                ("a_row_ptr", ctypes.c_void_p), ("a_col_idx", ctypes.c_void_p), ("a_values", ctypes.c_void_p),
                ("a_nnz", ctypes.c_uint64),
                ("b_row_ptr", ctypes.c_void_p), ("b_col_idx", ctypes.c_void_p), ("b_values", ctypes.c_void_p),
                ("b_nnz", ctypes.c_uint64),
                ("c_row_ptr", ctypes.c_void_p), ("c_col_idx", ctypes.c_void_p), ("c_values", ctypes.c_void_p),
                ("c_nnz", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    """nkr_g16_tuning: per-context choices, 0 = default (see include/nkr_g16.h)."""
    FIELDS = ("witness_window_bits", "h_window_bits", "ntt_lds_log", "ntt_radix_log", "ntt_split_policy",
              "msm_target_pieces_log", "witness_g1_piece_min")
    _fields_ = [(n, ctypes.c_uint32) for n in FIELDS]

    @classmethod
    def of(cls, d) -> "Tuning":
        unknown = set(d) - set(cls.FIELDS)
        if unknown:
            raise ValueError(f"unknown tuning fields {sorted(unknown)}")
        return cls(**{k: int(v) for k, v in d.items()})


class Pk(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "a_query", "b_g1_query", "b_g2_query", "l_query", "h_query",
        "alpha_g1", "beta_g1", "beta_g2", "delta_g1", "delta_g2")]


class PkOut(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "a_query", "b_g1_query", "b_g2_query", "l_query", "h_query",
        "alpha_g1", "beta_g1", "beta_g2", "delta_g1", "delta_g2")]


class VkOut(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("alpha_g1_beta_g2", "gamma_g2", "delta_g2", "ic")]


class Vk(ctypes.Structure):
    _fields_ = [("alpha_g1_beta_g2", ctypes.c_void_p), ("gamma_g2", ctypes.c_void_p),
                ("delta_g2", ctypes.c_void_p), ("ic", ctypes.c_void_p), ("ic_len", ctypes.c_uint64)]


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NkrError(f"{LIB_PATH} is not built (run __graft_entry__.build() or make -C nkr-zkp_amd)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int
        L.nkr_last_error.restype = ctypes.c_char_p
        L.nkr_version.restype = ctypes.c_char_p
        sigs = {
            "nkr_device_count": [ctypes.POINTER(ctypes.c_int)],
            "nkr_g16_ctx_create": [ctypes.POINTER(Circuit), ctypes.POINTER(Pk), i32, ctypes.POINTER(vp)],
            "nkr_g16_ctx_create_tuned": [ctypes.POINTER(Circuit), ctypes.POINTER(Pk), i32, ctypes.POINTER(Tuning),
                                         ctypes.POINTER(vp)],
            "nkr_g16_ctx_destroy": [vp],
            "nkr_g16_ctx_device_bytes": [vp, ctypes.POINTER(u64)],
            "nkr_g16_prove": [vp, vp, vp],
            "nkr_g16_prove_ex": [vp, vp, vp, vp, vp, vp, vp],
            "nkr_g16_blind": [ctypes.POINTER(Pk), vp, vp, vp, vp],
            "nkr_g16_prove_dev": [vp, vp, vp],
            "nkr_g16_check_witness": [ctypes.POINTER(Circuit), vp, ctypes.POINTER(ctypes.c_int)],
            "nkr_g16_prove_blind": [vp, vp, vp, vp, vp, vp],
            "nkr_g16_prove_blind_dev": [vp, vp, vp, vp, vp, vp],
            "nkr_g16_profile": [vp, i32],
            "nkr_g16_profile_read": [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64)],
            "nkr_dev_copy": [vp, vp, u64, vp],
            "nkr_valu_probe": [i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
            "nkr_ntt_set_shape": [u32, u32, u32],
            "nkr_msm_set_shape": [u32, u32],
            "nkr_g16_quotient": [vp, vp, vp, u64, vp, i32],
            "nkr_fr_ntt": [vp, u32, i32, i32],
            "nkr_fr_ntt_dev": [vp, u32, i32, vp],
            "nkr_fr_ntt_dev_noalloc": [vp, u32, i32, vp],
            "nkr_spmv": [u64, u64, vp, vp, vp, u64, vp, vp, i32],
            "nkr_spmv_create": [u64, u64, vp, vp, vp, u64, i32, ctypes.POINTER(vp)],
            "nkr_spmv_run_dev": [vp, vp, vp, vp],
            "nkr_spmv_destroy": [vp],
            "nkr_g1_msm_create": [vp, u64, u32, i32, ctypes.POINTER(vp)],
            "nkr_g2_msm_create": [vp, u64, u32, i32, ctypes.POINTER(vp)],
            "nkr_msm_run": [vp, vp, vp],
            "nkr_msm_run_dev": [vp, vp, vp, vp],
            "nkr_msm_window_bits": [vp, ctypes.POINTER(u32)],
            "nkr_msm_destroy": [vp],
            "nkr_g1_msm": [vp, vp, u64, vp, i32],
            "nkr_g2_msm": [vp, vp, u64, vp, i32],
            "nkr_field_op": [i32, vp, vp, vp, u64, i32],
            "nkr_g1_fixed_base_mul": [vp, vp, u64, vp, i32],
            "nkr_g2_fixed_base_mul": [vp, vp, u64, vp, i32],
            "nkr_g1_fixed_base_mul_dev": [vp, vp, u64, vp, vp],
            "nkr_g2_fixed_base_mul_dev": [vp, vp, u64, vp, vp],
            "nkr_dev_alloc": [u64, ctypes.POINTER(vp)],
            "nkr_dev_free": [vp],
            "nkr_memcpy_htod": [vp, vp, u64],
            "nkr_memcpy_dtoh": [vp, vp, u64],
            "nkr_device_synchronize": [],
            "nkr_device_pci_bus_id": [i32, ctypes.c_char_p, i32],
            "nkr_g16_verify": [ctypes.POINTER(Vk), vp, u64, vp, ctypes.POINTER(ctypes.c_int)],
            "nkr_g16_setup": [ctypes.POINTER(Circuit), vp, i32, ctypes.POINTER(PkOut), ctypes.POINTER(VkOut)],
            "nkr_g16_batch_create": [i32, ctypes.POINTER(vp)],
            "nkr_g16_batch_prove": [vp, ctypes.POINTER(BatchItem), u64],
            "nkr_g16_batch_destroy": [vp],
            "nkr_g16_batch_set_graphs": [vp, i32],
            "nkr_g16_server_warm": [vp, ctypes.POINTER(BatchItem), u64],
            "nkr_g16_batch_graph_active": [vp, ctypes.POINTER(i32)],
            "nkr_g16_server_create": [i32, u32, u32, u64, ctypes.POINTER(vp)],
            "nkr_g16_server_submit": [vp, vp, vp, vp, vp, vp, vp, ctypes.POINTER(u64)],
            "nkr_g16_server_wait": [vp, u64, ctypes.POINTER(ctypes.c_double)],
            "nkr_g16_server_request_bytes": [vp, ctypes.POINTER(u64)],
            "nkr_g16_server_stats": [vp, ctypes.POINTER(u64)],
            "nkr_g16_server_memory": [vp, ctypes.POINTER(u64)],
            "nkr_g16_server_destroy": [vp],
            "nkr_bn254_pairing": [vp, vp, vp],
            "nkr_bn254_pairing_product": [vp, vp, u64, vp],
            "nkr_bn254_miller_loop": [vp, vp, vp],
            "nkr_bn254_final_exponentiation": [vp, vp],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def header_functions() -> list[str]:
    """Names of every function declared in include/nkr_g16.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(nkr_\w+)\s*\(", text, re.M)))


def _check(rc: int) -> None:
    if rc != 0:
        raise NkrError(lib().nkr_last_error().decode())


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _words(a, width: int) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64).reshape(-1, width))


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(lib().nkr_device_count(ctypes.byref(n)))
    return n.value


def device_pci_bus_id(device: int = 0) -> str:
    """PCI address of `device`, lower case ("0000:75:00.0")."""
    buf = ctypes.create_string_buffer(64)
    _check(lib().nkr_device_pci_bus_id(device, buf, len(buf)))
    return buf.value.decode().lower()


FIELD_OPS = {"fr_mul": 0, "fr_add": 1, "fr_sub": 2, "fq_mul": 3, "fq_add": 4, "fq_sub": 5,
             "fq2_mul": 6, "fq2_sqr": 7, "fr_inv": 8}


def field_op(name: str, a, b=None, device: int = 0) -> np.ndarray:
    op = FIELD_OPS[name]
    width = 8 if name.startswith("fq2") else 4
    a = _words(a, width)
    b = _words(b, width) if b is not None else None
    out = np.zeros_like(a)
    _check(lib().nkr_field_op(op, _p(a), _p(b) if b is not None else None, _p(out), a.shape[0], device))
    return out


FFT, IFFT, COSET_FFT, ICOSET_FFT = 0, 1, 2, 3


def ntt(a, kind: int, device: int = 0) -> np.ndarray:
    out = _words(a, 4).copy()
    n = out.shape[0]
    log_n = n.bit_length() - 1
    if n != 1 << log_n:
        raise ValueError("size must be a power of two")
    _check(lib().nkr_fr_ntt(_p(out), log_n, kind, device))
    return out


def quotient(az, bz, cz, device: int = 0) -> np.ndarray:
    az, bz, cz = (_words(v, 4) for v in (az, bz, cz))
    h = np.zeros_like(az)
    _check(lib().nkr_g16_quotient(_p(az), _p(bz), _p(cz), az.shape[0], _p(h), device))
    return h


def spmv(rows, cols, row_ptr, col_idx, values, z, device: int = 0) -> np.ndarray:
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.uint32)
    values = _words(values, 4)
    z = _words(z, 4)
    y = np.zeros((rows, 4), dtype=np.uint64)
    _check(lib().nkr_spmv(rows, cols, _p(row_ptr), _p(col_idx) if col_idx.size else None,
                          _p(values) if values.size else None, values.shape[0], _p(z), _p(y), device))
    return y


class Spmv:
    """One CSR matrix uploaded once (spMVMContext), multiplied with device vectors."""

    def __init__(self, rows, cols, row_ptr, col_idx, values, device: int = 0):
        row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        col_idx = np.ascontiguousarray(col_idx, dtype=np.uint32)
        values = _words(values, 4)
        self.rows, self.cols = rows, cols
        h = ctypes.c_void_p()
        _check(lib().nkr_spmv_create(rows, cols, _p(row_ptr), _p(col_idx) if col_idx.size else None,
                                     _p(values) if values.size else None, values.shape[0], device,
                                     ctypes.byref(h)))
        self.h = h

    def run_dev(self, z_dev: int, y_dev: int, stream: int | None = None) -> None:
        _check(lib().nkr_spmv_run_dev(self.h, z_dev, y_dev, stream))

    def close(self) -> None:
        if self.h:
            lib().nkr_spmv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Msm:
    """Prepared-bases MSM (MSMContext: load_bases once, run per scalar vector)."""

    def __init__(self, bases, g2: bool = False, c: int = 0, device: int = 0):
        self.g2 = g2
        width = 24 if g2 else 12
        b = _words(bases, width)
        self.n = b.shape[0]
        h = ctypes.c_void_p()
        fn = lib().nkr_g2_msm_create if g2 else lib().nkr_g1_msm_create
        _check(fn(_p(b), self.n, c, device, ctypes.byref(h)))
        self.h = h

    @property
    def window_bits(self) -> int:
        c = ctypes.c_uint32()
        _check(lib().nkr_msm_window_bits(self.h, ctypes.byref(c)))
        return c.value

    def run(self, scalars) -> np.ndarray:
        s = _words(scalars, 4)
        assert s.shape[0] == self.n
        out = np.zeros(24 if self.g2 else 12, dtype=np.uint64)
        _check(lib().nkr_msm_run(self.h, _p(s), _p(out)))
        return out

    def run_dev(self, scalars_dev: int, stream: int | None = None) -> np.ndarray:
        out = np.zeros(24 if self.g2 else 12, dtype=np.uint64)
        _check(lib().nkr_msm_run_dev(self.h, scalars_dev, _p(out), stream))
        return out

    def close(self) -> None:
        if self.h:
            lib().nkr_msm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fixed_base_mul(base, scalars, g2: bool = False, device: int = 0) -> np.ndarray:
    """out[i] = scalars[i] * base (normalised points)."""
    width = 24 if g2 else 12
    b = np.ascontiguousarray(np.asarray(base, dtype=np.uint64).reshape(width))
    s = _words(scalars, 4)
    out = np.zeros((s.shape[0], width), dtype=np.uint64)
    fn = lib().nkr_g2_fixed_base_mul if g2 else lib().nkr_g1_fixed_base_mul
    _check(fn(_p(b), _p(s), s.shape[0], _p(out), device))
    return out


def msm(bases, scalars, g2: bool = False, device: int = 0) -> np.ndarray:
    width = 24 if g2 else 12
    b = _words(bases, width)
    s = _words(scalars, 4)
    out = np.zeros(width, dtype=np.uint64)
    fn = lib().nkr_g2_msm if g2 else lib().nkr_g1_msm
    _check(fn(_p(b), _p(s), b.shape[0], _p(out), device))
    return out


def _circuit_struct(circ, keep: list) -> Circuit:
    """Circuit view of an object with rows, cols, public_inputs and a, b, c =
    (row_ptr, col_idx, values); the arrays it points at are appended to keep."""
    def arr(x, dtype, width=None):
        a = np.ascontiguousarray(np.asarray(x, dtype=dtype))
        if width:
            a = a.reshape(-1, width)
        keep.append(a)
        return a

    mats = []
    for m in (circ.a, circ.b, circ.c):
        mats.append((arr(m[0], np.uint32), arr(m[1], np.uint32), arr(m[2], np.uint64, 4)))
    return Circuit(circ.rows, circ.cols, circ.public_inputs,
                   *[v for (rp, ci, va) in mats for v in (_p(rp), _p(ci) if ci.size else None,
                                                           _p(va) if va.size else None, va.shape[0])])


class Keys:
    """Proving and verification key arrays (uint64 words, reference layouts)."""


def setup(circ, trapdoors, device: int = 0) -> Keys:
    """Backend::setup with caller-supplied trapdoors (t, alpha, beta, gamma,
    delta as a (5, 4) Montgomery array): one object holding the proving key
    fields (a_query, ..., delta_g2) and the verification key fields
    (alpha_g1_beta_g2, gamma_g2, ic; delta_g2 is shared)."""
    keep = []
    c = _circuit_struct(circ, keep)
    td = _words(trapdoors, 4)
    assert td.shape[0] == 5
    k = Keys()
    rows, cols, pub = circ.rows, circ.cols, circ.public_inputs
    shapes = {"a_query": (cols, 12), "b_g1_query": (cols, 12), "b_g2_query": (cols, 24), "l_query": (cols, 12),
              "h_query": (rows, 12), "alpha_g1": (12,), "beta_g1": (12,), "beta_g2": (24,), "delta_g1": (12,),
              "delta_g2": (24,), "alpha_g1_beta_g2": (GT_WORDS,), "gamma_g2": (24,), "ic": (pub + 1, 12)}
    for name, shape in shapes.items():
        setattr(k, name, np.zeros(shape, dtype=np.uint64))
    pk_out = PkOut(*[_p(getattr(k, n)) for n, _ in PkOut._fields_])
    vk_out = VkOut(_p(k.alpha_g1_beta_g2), _p(k.gamma_g2), _p(k.delta_g2), _p(k.ic))
    _check(lib().nkr_g16_setup(ctypes.byref(c), _p(td), device, ctypes.byref(pk_out), ctypes.byref(vk_out)))
    return k


class BatchItem(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("witness", ctypes.c_void_p), ("r", ctypes.c_void_p),
                ("s", ctypes.c_void_p), ("raw_out", ctypes.c_void_p), ("proof_out", ctypes.c_void_p)]


class Batch:
    """nkr_g16_batch: many proofs (any mix of Prover contexts) as one launch sequence."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().nkr_g16_batch_create(device, ctypes.byref(h)))
        self.h = h

    def prove(self, items):
        """items: (Prover, witness, r, s) tuples -> list of (raw words, proof words)."""
        keep, outs = [], []
        arr = (BatchItem * max(len(items), 1))()
        for i, (pv, w, r, s) in enumerate(items):
            w = _words(w, 4)
            assert w.shape[0] == pv.cols
            r = np.ascontiguousarray(r, dtype=np.uint64).reshape(4)
            s = np.ascontiguousarray(s, dtype=np.uint64).reshape(4)
            raw = np.zeros(72, dtype=np.uint64)
            proof = np.zeros(48, dtype=np.uint64)
            keep += [w, r, s]
            outs.append((raw, proof))
            arr[i] = BatchItem(pv.h.value, _p(w), _p(r), _p(s), _p(raw), _p(proof))
        _check(lib().nkr_g16_batch_prove(self.h, arr, len(items)))
        return outs

    def set_graphs(self, enabled: bool) -> None:
        """Capture a repeated composition into a HIP graph and replay it (default on)."""
        _check(lib().nkr_g16_batch_set_graphs(self.h, 1 if enabled else 0))

    def graph_active(self) -> bool:
        v = ctypes.c_int32()
        _check(lib().nkr_g16_batch_graph_active(self.h, ctypes.byref(v)))
        return bool(v.value)

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().nkr_g16_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """nkr_g16_server: continuous-batching request loop over resident contexts."""

    def __init__(self, device: int = 0, workers: int = 2, max_batch: int = 64, budget_bytes: int = 64 << 30):
        h = ctypes.c_void_p()
        _check(lib().nkr_g16_server_create(device, workers, max_batch, budget_bytes, ctypes.byref(h)))
        self.h = h
        self._live = {}

    def submit(self, pv, witness, r, s) -> int:
        w = _words(witness, 4)
        assert w.shape[0] == pv.cols
        r = np.ascontiguousarray(r, dtype=np.uint64).reshape(4)
        s = np.ascontiguousarray(s, dtype=np.uint64).reshape(4)
        raw = np.zeros(72, dtype=np.uint64)
        proof = np.zeros(48, dtype=np.uint64)
        t = ctypes.c_uint64()
        _check(lib().nkr_g16_server_submit(self.h, pv.h, _p(w), _p(r), _p(s), _p(raw), _p(proof), ctypes.byref(t)))
        self._live[t.value] = (w, r, s, raw, proof)
        return t.value

    def wait(self, ticket: int):
        """-> (raw words, proof words, latency ms)"""
        ms = ctypes.c_double()
        try:
            _check(lib().nkr_g16_server_wait(self.h, ticket, ctypes.byref(ms)))
        finally:
            w, r, s, raw, proof = self._live.pop(ticket)
        return raw, proof, ms.value

    def warm(self, items):
        """Every worker proves items ((Prover, witness, r, s) tuples) once, as
        one batch, so its buffers are sized before traffic; returns the
        outputs as Batch.prove does."""
        keep, outs = [], []
        arr = (BatchItem * max(len(items), 1))()
        for i, (pv, w, r, s) in enumerate(items):
            w = _words(w, 4)
            assert w.shape[0] == pv.cols
            r = np.ascontiguousarray(r, dtype=np.uint64).reshape(4)
            s = np.ascontiguousarray(s, dtype=np.uint64).reshape(4)
            raw = np.zeros(72, dtype=np.uint64)
            proof = np.zeros(48, dtype=np.uint64)
            keep += [w, r, s]
            outs.append((raw, proof))
            arr[i] = BatchItem(pv.h.value, _p(w), _p(r), _p(s), _p(raw), _p(proof))
        _check(lib().nkr_g16_server_warm(self.h, arr, len(items)))
        return outs

    def stats(self) -> dict:
        st = (ctypes.c_uint64 * 4)()
        _check(lib().nkr_g16_server_stats(self.h, st))
        return {"batches": st[0], "proofs": st[1], "max_batch": st[2], "peak_work_bytes": st[3]}

    def memory(self) -> dict:
        m = (ctypes.c_uint64 * 4)()
        _check(lib().nkr_g16_server_memory(self.h, m))
        return {"budget": m[0], "share": m[1], "held": m[2], "peak_held": m[3]}

    @staticmethod
    def request_bytes(pv) -> int:
        b = ctypes.c_uint64()
        _check(lib().nkr_g16_server_request_bytes(pv.h, ctypes.byref(b)))
        return b.value

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().nkr_g16_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Prover:
    """Groth16 backend context: prepare in the constructor, prove per witness."""

    def __init__(self, circ, pk, device: int = 0, tuning=None):
        """tuning: optional dict of nkr_g16_tuning fields (autotune.py makes them)."""
        self._keep = []
        self.rows, self.cols = circ.rows, circ.cols
        c = _circuit_struct(circ, self._keep)
        self.pk_struct = self._pk(pk)
        h = ctypes.c_void_p()
        if tuning:
            t = Tuning.of(tuning)
            _check(lib().nkr_g16_ctx_create_tuned(ctypes.byref(c), ctypes.byref(self.pk_struct), device,
                                                  ctypes.byref(t), ctypes.byref(h)))
        else:
            _check(lib().nkr_g16_ctx_create(ctypes.byref(c), ctypes.byref(self.pk_struct), device, ctypes.byref(h)))
        self.h = h

    def _pk(self, pk) -> Pk:
        """The proving key view; alpha/beta/delta may be absent (None): such a
        context proves but cannot blind."""
        arrs = {}
        for k in ("a_query", "b_g1_query", "b_g2_query", "l_query", "h_query",
                  "alpha_g1", "beta_g1", "beta_g2", "delta_g1", "delta_g2"):
            v = getattr(pk, k, None)
            if v is None and k.startswith(("alpha", "beta", "delta")):
                arrs[k] = None
                continue
            a = np.ascontiguousarray(np.asarray(v, dtype=np.uint64))
            self._keep.append(a)
            arrs[k] = a.ctypes.data
        return Pk(**arrs)

    def device_bytes(self) -> int:
        b = ctypes.c_uint64()
        _check(lib().nkr_g16_ctx_device_bytes(self.h, ctypes.byref(b)))
        return b.value

    def prove(self, witness, intermediates: bool = False):
        w = _words(witness, 4)
        assert w.shape[0] == self.cols
        raw = np.zeros(72, dtype=np.uint64)
        if not intermediates:
            _check(lib().nkr_g16_prove(self.h, _p(w), _p(raw)))
            return raw
        bufs = [np.zeros((self.rows, 4), dtype=np.uint64) for _ in range(4)]
        _check(lib().nkr_g16_prove_ex(self.h, _p(w), _p(raw), *[_p(b) for b in bufs]))
        return raw, tuple(bufs)

    def prove_blind(self, witness, r, s):
        """prove + blind_proof in one call; returns (raw words, proof words)."""
        w = _words(witness, 4)
        assert w.shape[0] == self.cols
        r = np.ascontiguousarray(r, dtype=np.uint64).reshape(4)
        s = np.ascontiguousarray(s, dtype=np.uint64).reshape(4)
        raw = np.zeros(72, dtype=np.uint64)
        proof = np.zeros(48, dtype=np.uint64)
        _check(lib().nkr_g16_prove_blind(self.h, _p(w), _p(r), _p(s), _p(raw), _p(proof)))
        return raw, proof

    def blind(self, raw, r, s) -> np.ndarray:
        raw = np.ascontiguousarray(raw, dtype=np.uint64)
        r = np.ascontiguousarray(r, dtype=np.uint64).reshape(4)
        s = np.ascontiguousarray(s, dtype=np.uint64).reshape(4)
        out = np.zeros(48, dtype=np.uint64)
        _check(lib().nkr_g16_blind(ctypes.byref(self.pk_struct), _p(raw), _p(r), _p(s), _p(out)))
        return out

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().nkr_g16_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def blind(pk, raw, r, s) -> np.ndarray:
    keep = []
    arrs = {}
    for k in ("a_query", "b_g1_query", "b_g2_query", "l_query", "h_query",
              "alpha_g1", "beta_g1", "beta_g2", "delta_g1", "delta_g2"):
        a = np.ascontiguousarray(np.asarray(getattr(pk, k), dtype=np.uint64))
        keep.append(a)
        arrs[k] = a.ctypes.data
    p = Pk(**arrs)
    raw = np.ascontiguousarray(raw, dtype=np.uint64)
    r = np.ascontiguousarray(r, dtype=np.uint64).reshape(4)
    s = np.ascontiguousarray(s, dtype=np.uint64).reshape(4)
    out = np.zeros(48, dtype=np.uint64)
    _check(lib().nkr_g16_blind(ctypes.byref(p), _p(raw), _p(r), _p(s), _p(out)))
    return out


# ------------------------------------------------------------ verifier (host)
GT_WORDS = 48  # Fq12: 12 Fq elements of 4 words


def pairing_product(g1s, g2s) -> np.ndarray:
    """prod_i e(P_i, Q_i): g1s (k, 12) words, g2s (k, 24) words -> GT (48 words)."""
    g1s = _words(g1s, 12)
    g2s = _words(g2s, 24)
    assert len(g1s) == len(g2s)
    out = np.zeros(GT_WORDS, dtype=np.uint64)
    _check(lib().nkr_bn254_pairing_product(_p(g1s), _p(g2s), len(g1s), _p(out)))
    return out


def pairing(g1, g2) -> np.ndarray:
    g1 = _words(g1, 12)
    g2 = _words(g2, 24)
    out = np.zeros(GT_WORDS, dtype=np.uint64)
    _check(lib().nkr_bn254_pairing(_p(g1), _p(g2), _p(out)))
    return out


def miller_loop(g1, g2) -> np.ndarray:
    g1 = _words(g1, 12)
    g2 = _words(g2, 24)
    out = np.zeros(GT_WORDS, dtype=np.uint64)
    _check(lib().nkr_bn254_miller_loop(_p(g1), _p(g2), _p(out)))
    return out


def final_exponentiation(f) -> np.ndarray:
    f = np.ascontiguousarray(f, dtype=np.uint64).reshape(GT_WORDS)
    out = np.zeros(GT_WORDS, dtype=np.uint64)
    _check(lib().nkr_bn254_final_exponentiation(_p(f), _p(out)))
    return out


def verify(vk, public_input, proof) -> bool:
    """Groth16 verify (protocol.cpp:96-108).  vk: object with alpha_g1_beta_g2
    (48 words), gamma_g2, delta_g2 (24 words) and ic ((n+1, 12) words)."""
    ab = np.ascontiguousarray(vk.alpha_g1_beta_g2, dtype=np.uint64).reshape(GT_WORDS)
    gamma = _words(vk.gamma_g2, 24)
    delta = _words(vk.delta_g2, 24)
    ic = _words(vk.ic, 12)
    x = _words(public_input, 4) if len(public_input) else np.zeros((0, 4), dtype=np.uint64)
    proof = np.ascontiguousarray(proof, dtype=np.uint64).reshape(48)
    v = Vk(_p(ab), _p(gamma), _p(delta), _p(ic), len(ic))
    ok = ctypes.c_int(0)
    _check(lib().nkr_g16_verify(ctypes.byref(v), _p(x), len(x), _p(proof), ctypes.byref(ok)))
    return bool(ok.value)
