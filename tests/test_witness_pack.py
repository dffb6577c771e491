"""The witness upload's packing rule (nkr-zkp_amd/csrc/witness_pack.h),
compiled for the CPU: the masks mark exactly the elements equal to 0, to 1
(Montgomery form 2^256 mod r) and the others; the others are packed in order
from each range's first index; and expanding the packed form as the device
kernel does (a value's index is its group's offset plus the number of marked
elements below it in the group) gives the witness back, for sizes that are
not multiples of a group or a range, all-zero, all-one and dense witnesses."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ONE = np.array(O.int_to_words(O.to_mont_int(1, O.R)), dtype=np.uint64)


@pytest.fixture(scope="module")
def pack_bin(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("needs g++")
    out = str(tmp_path_factory.mktemp("pack") / "witness_pack_check")
    subprocess.run([cxx, "-O2", "-std=c++17", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "nkr-zkp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "host", "witness_pack_check.cpp"), "-o", out], check=True)
    return out


def pack(binary, w, chunk, tmp_path):
    n = w.shape[0]
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    np.ascontiguousarray(w, dtype=np.uint64).tofile(src)
    subprocess.run([binary, str(src), str(n), str(chunk), str(dst)], check=True)
    raw = np.fromfile(dst, dtype=np.uint8)
    groups = (n + 63) // 64
    ranges = (n + chunk - 1) // chunk
    at = n * 32
    vals = raw[:at].view(np.uint64).reshape(n, 4)
    masks = raw[at:at + 16 * groups].view(np.uint64).reshape(groups, 2)
    at += 16 * groups
    offsets = raw[at:at + 4 * groups].view(np.uint32)
    at += 4 * groups
    counts = raw[at:at + 4 * ranges].view(np.uint32)
    return vals, masks, offsets, counts


def expand(vals, masks, offsets, n):
    """The device kernel's rule (k_expand_witness), element by element."""
    out = np.zeros((n, 4), dtype=np.uint64)
    for i in range(n):
        g, bit = i >> 6, 1 << (i & 63)
        other, one = int(masks[g, 0]), int(masks[g, 1])
        if other & bit:
            out[i] = vals[int(offsets[g]) + bin(other & (bit - 1)).count("1")]
        elif one & bit:
            out[i] = ONE
    return out


def witness(n, seed, p_zero, p_one):
    w = O.fr_random(seed, n)
    rng = np.random.default_rng(seed)
    u = rng.random(n)
    w[u < p_zero] = 0
    w[(u >= p_zero) & (u < p_zero + p_one)] = ONE
    return w


@pytest.mark.parametrize("n,chunk,p_zero,p_one", [(1, 64, 0.0, 0.0), (63, 64, 0.5, 0.25), (64, 64, 1.0, 0.0),
                                                  (65, 128, 0.0, 1.0), (1000, 256, 0.45, 0.45),
                                                  (5000, 1024, 0.9, 0.05), (4096, 4096, 0.0, 0.0),
                                                  (3333, 64, 0.3, 0.3)])
def test_pack_round_trip(pack_bin, tmp_path, n, chunk, p_zero, p_one):
    w = witness(n, n + chunk, p_zero, p_one)
    vals, masks, offsets, counts = pack(pack_bin, w, chunk, tmp_path)
    is_zero = ~w.any(axis=1)
    is_one = (w == ONE).all(axis=1)
    other = ~(is_zero | is_one)
    for g in range((n + 63) // 64):
        sl = slice(64 * g, min(n, 64 * g + 64))
        assert int(masks[g, 0]) == sum(1 << k for k, x in enumerate(other[sl]) if x)
        assert int(masks[g, 1]) == sum(1 << k for k, x in enumerate(is_one[sl]) if x)
    for c in range(len(counts)):
        sl = slice(c * chunk, min(n, (c + 1) * chunk))
        assert counts[c] == other[sl].sum()
        assert np.array_equal(vals[c * chunk:c * chunk + counts[c]], w[sl][other[sl]])
    assert np.array_equal(expand(vals, masks, offsets, n), w)


@pytest.mark.parametrize("sizes,chunk", [([1, 1], 64), ([63, 65, 64, 1], 64), ([1000, 5000, 3], 1024),
                                         ([4096, 129, 4096], 4096), ([200, 7, 333], 128)])
def test_pack_several_witnesses(pack_bin, tmp_path, sizes, chunk):
    """The batch upload's layout: witnesses end to end from multiples of 64,
    each packed from its own buffer; expanding the shared space gives every
    witness back at its base, and the padding between them expands to zero."""
    ws = [witness(z, 7 * z + i, 0.4, 0.3) for i, z in enumerate(sizes)]
    bases, at = [], 0
    for z in sizes:
        bases.append(at)
        at = (at + z + 63) // 64 * 64
    n = at
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    np.ascontiguousarray(np.concatenate(ws), dtype=np.uint64).tofile(src)
    subprocess.run([pack_bin, str(src), str(n), str(chunk), str(dst), ",".join(map(str, sizes))], check=True)
    raw = np.fromfile(dst, dtype=np.uint8)
    groups = n // 64
    vals = raw[:n * 32].view(np.uint64).reshape(n, 4)
    masks = raw[n * 32:n * 32 + 16 * groups].view(np.uint64).reshape(groups, 2)
    offsets = raw[n * 32 + 16 * groups:n * 32 + 20 * groups].view(np.uint32)
    counts = raw[n * 32 + 20 * groups:].view(np.uint32)
    assert len(counts) == sum((z + chunk - 1) // chunk for z in sizes)
    out = expand(vals, masks, offsets, n)
    covered = np.zeros(n, dtype=bool)
    for w, b, z in zip(ws, bases, sizes):
        assert np.array_equal(out[b:b + z], w)
        covered[b:b + z] = True
    assert not out[~covered].any()
