"""nkr-g16-verify, the groth16-verify driver (protocols/groth16/app/verify.cpp:
26-77): reads r1cs, witness (public inputs), verification key and proof files
and runs the pairing check on the host.  The GPU case proves with
nkr-g16-prove and verifies the file it wrote."""
import os
import subprocess

import numpy as np
import pytest

import archive
import goldens

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "nkr-zkp_amd", "bin")
VERIFY = os.path.join(BIN, "nkr-g16-verify")
PROVE = os.path.join(BIN, "nkr-g16-prove")


class Vk:
    pass


def stage(tmp_path, proof=None, witness=None):
    e = goldens.E2E()
    v = goldens.load("bn254_pairing.json")["matmul2_vk"]
    vk = Vk()
    vk.alpha_g1_beta_g2 = goldens.fe(v["alpha_g1_beta_g2"]).reshape(-1)
    vk.gamma_g2 = goldens.point(v["gamma_g2"])
    vk.delta_g2 = goldens.point(v["delta_g2"])
    vk.ic = goldens.points(e.d["vk_ic"])
    archive.write_circuit(tmp_path / "m.bin", e)
    archive.write_witness(tmp_path / "m.witness.bin", e.witness if witness is None else witness)
    archive.write_proving_key(tmp_path / "m.pk.bin", e)
    archive.write_verification_key(tmp_path / "m.vk.bin", vk)
    if proof is not False:
        archive.write_proof(tmp_path / "m.proof.bin", e.proof if proof is None else proof)
    return e


def verify(tmp_path, *extra):
    return subprocess.run([VERIFY, "--path", str(tmp_path), *extra, "m"], capture_output=True, text=True,
                          timeout=120)


def test_accepts_golden_proof(tmp_path):
    stage(tmp_path)
    r = verify(tmp_path, "--profile")
    assert r.returncode == 0, r.stderr
    assert "Proof verification succeeded." in r.stdout
    assert "[profile] groth16.verify.pairing_check" in r.stdout


def test_rejects_tampered_proof(tmp_path):
    e = goldens.E2E()
    bad = e.proof.copy()
    bad[36:] = e.proof[:12]  # C := A
    stage(tmp_path, proof=bad)
    r = verify(tmp_path)
    assert r.returncode == 1 and "Proof verification FAILED." in r.stdout


def test_rejects_other_public_input(tmp_path):
    e = goldens.E2E()
    w = e.witness.copy()
    w[[1, 2]] = w[[2, 1]]
    assert not np.array_equal(w, e.witness)
    stage(tmp_path, witness=w)
    r = verify(tmp_path)
    assert r.returncode == 1 and "FAILED" in r.stdout


def test_short_witness_and_missing_file(tmp_path):
    e = goldens.E2E()
    stage(tmp_path, witness=e.witness[:3])
    r = verify(tmp_path)
    assert r.returncode == 1 and "Witness is too short" in r.stderr
    r = verify(tmp_path, "--proof", str(tmp_path / "absent.bin"))
    assert r.returncode == 1 and r.stderr.startswith("verify: ")


def test_usage():
    r = subprocess.run([VERIFY], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2 and "Usage" in r.stderr


@pytest.mark.gpu
def test_prove_then_verify(tmp_path):
    stage(tmp_path, proof=False)
    r = subprocess.run([PROVE, "--path", str(tmp_path), "--seed", "77", "m"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    r = verify(tmp_path)
    assert r.returncode == 0 and "Proof verification succeeded." in r.stdout, r.stdout + r.stderr
