"""CPU pin of the identity behind the two-stream inner product (k1w_*
kernels): with IPGK0 = sigma(a_i) - 2^26 IPGK1 (the digits of the key switch,
base_pir.rs:544-579) and the chain a_(i+1) = sigma(a_i) gka0 + IPGK1_i (gka1 -
2^26 gka0) + qry_a TE_i (base_pir.rs:838-861, a_0 = 0), the three-array
main_pir_loop sum (base_pir.rs:1153-1208) equals a two-array sum with
per-query weights. two_stream_check.c restates the device formula in the
output-major Galois coordinates and compares it with the oracle's K1 for
every output, modulus and lane, at ring sizes small enough for brute force;
a weight index planted off by one must fail."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    out = str(tmp_path_factory.mktemp("two_stream") / "two_stream_check")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-w", "-o", out, os.path.join(HERE, "two_stream_check.c"), "-lm"],
                   check=True)
    return out


@pytest.mark.parametrize("logN", [5, 7, 8])
def test_two_stream_identity_matches_oracle_k1(checker, logN):
    r = subprocess.run([checker, str(logN), "0"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(": 0 mismatches") == 2


@pytest.mark.parametrize("shift", [1, -1])
def test_two_stream_identity_detects_a_wrong_weight_index(checker, shift):
    r = subprocess.run([checker, "5", str(shift)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1
    assert ": 0 mismatches" not in r.stdout
