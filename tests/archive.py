"""A writer/reader of the reference's BinaryArchive artifact format (magic
"LKZ1", u32 version 1, trivially-copyable values, u64-byte-length-prefixed
vectors; common/include/binary_archive.hpp:17-97, protocols/groth16/common/src/io.cpp)
for tests, independent of the C++ implementation under test."""
import struct

import numpy as np

MAGIC = b"LKZ1"


def _vec(a) -> bytes:
    b = np.ascontiguousarray(a).tobytes()
    return struct.pack("<Q", len(b)) + b


def write_circuit(path, circ):
    out = MAGIC + struct.pack("<I", 1) + struct.pack("<QQ", circ.public_inputs, circ.cols)
    for m in (circ.a, circ.b, circ.c):
        out += _vec(np.asarray(m[0], dtype=np.uint32)) + _vec(np.asarray(m[1], dtype=np.uint32)) + \
            _vec(np.asarray(m[2], dtype=np.uint64))
    open(path, "wb").write(out)


def write_witness(path, w):
    open(path, "wb").write(MAGIC + struct.pack("<I", 1) + _vec(np.asarray(w, dtype=np.uint64)))


def write_proving_key(path, pk):
    out = MAGIC + struct.pack("<I", 1)
    for k in ("a_query", "b_g1_query", "b_g2_query", "l_query", "h_query"):
        out += _vec(np.asarray(getattr(pk, k), dtype=np.uint64))
    for k in ("alpha_g1", "beta_g1", "beta_g2", "delta_g1", "delta_g2"):
        out += np.ascontiguousarray(np.asarray(getattr(pk, k), dtype=np.uint64)).tobytes()
    open(path, "wb").write(out)


def write_verification_key(path, vk):
    """alpha_g1_beta_g2 (GT), gamma_g2, delta_g2, then the ic vector (io.cpp:79-84)."""
    out = MAGIC + struct.pack("<I", 1)
    for k in ("alpha_g1_beta_g2", "gamma_g2", "delta_g2"):
        out += np.ascontiguousarray(np.asarray(getattr(vk, k), dtype=np.uint64)).tobytes()
    out += _vec(np.asarray(vk.ic, dtype=np.uint64))
    open(path, "wb").write(out)


def write_proof(path, proof):
    open(path, "wb").write(MAGIC + struct.pack("<I", 1) +
                           np.ascontiguousarray(np.asarray(proof, dtype=np.uint64)).tobytes())


def read_proof(path) -> np.ndarray:
    data = open(path, "rb").read()
    assert data[:4] == MAGIC and struct.unpack("<I", data[4:8])[0] == 1
    body = np.frombuffer(data[8:], dtype=np.uint64)
    assert body.size == 48
    return body.copy()
