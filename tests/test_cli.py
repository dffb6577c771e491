"""nkr-g16-prove, the groth16-prove driver over the MI355X backend
(protocols/groth16/app/prove.cpp:27-99): reads the reference's artifact files,
validates, proves, blinds with seeded r and s, writes the proof file.
Failure handling runs on the CPU; the full prove needs the GPU."""
import os
import subprocess

import numpy as np
import pytest

import archive
import goldens
import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "nkr-zkp_amd", "bin", "nkr-g16-prove")


def stage(tmp_path, e, witness=None, circ=None):
    archive.write_circuit(tmp_path / "m.bin", circ or e)
    archive.write_witness(tmp_path / "m.witness.bin", e.witness if witness is None else witness)
    archive.write_proving_key(tmp_path / "m.pk.bin", e)


def run(tmp_path, *extra):
    return subprocess.run([CLI, "--path", str(tmp_path), *extra, "m"], capture_output=True, text=True, timeout=120)


def test_rejects_truncated_archive(tmp_path):
    e = goldens.E2E()
    stage(tmp_path, e)
    data = open(tmp_path / "m.pk.bin", "rb").read()
    open(tmp_path / "m.pk.bin", "wb").write(data[: len(data) // 2])
    r = run(tmp_path)
    assert r.returncode == 1 and "BinaryArchive" in r.stderr


def test_rejects_bad_magic(tmp_path):
    e = goldens.E2E()
    stage(tmp_path, e)
    data = bytearray(open(tmp_path / "m.bin", "rb").read())
    data[0:4] = b"XXXX"
    open(tmp_path / "m.bin", "wb").write(bytes(data))
    r = run(tmp_path)
    assert r.returncode == 1 and "incompatible or truncated" in r.stderr


def test_rejects_unsatisfying_witness(tmp_path):
    e = goldens.E2E()
    w = e.witness.copy()
    w[5] = O.int_to_words(O.to_mont_int(12345, O.R))
    stage(tmp_path, e, witness=w)
    r = run(tmp_path)
    assert r.returncode == 1 and "witness does not satisfy the R1CS" in r.stderr


def test_rejects_out_of_range_column(tmp_path):
    e = goldens.E2E()

    class C:
        pass
    c = C()
    c.public_inputs, c.cols, c.rows = e.public_inputs, e.cols, e.rows
    c.a, c.b = e.a, e.b
    ci = e.c[1].copy()
    ci[0] = e.cols + 3
    c.c = (e.c[0], ci, e.c[2])
    stage(tmp_path, e, circ=c)
    r = run(tmp_path)
    assert r.returncode == 1 and "out-of-range column" in r.stderr


def test_usage():
    r = subprocess.run([CLI, "--bogus"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2 and "Usage" in r.stderr


@pytest.mark.gpu
def test_cli_proof_matches_oracle(tmp_path):
    e = goldens.E2E()
    stage(tmp_path, e)
    r = run(tmp_path, "--seed", "1234", "--profile")
    assert r.returncode == 0, r.stderr
    assert "Groth16 prove backend: hip" in r.stdout
    assert "[profile] groth16.prove.total" in r.stdout
    proof = archive.read_proof(tmp_path / "m.proof.bin")
    rs = O.fr_random(1234, 2)
    raw = O.prove(e, e, e.witness)
    want = O.blind(e, raw, rs[0], rs[1])
    aff = np.concatenate([O.g1_to_affine(want[:12])[0], O.g2_to_affine(want[12:36])[0], O.g1_to_affine(want[36:])[0]])
    assert np.array_equal(proof, aff)
