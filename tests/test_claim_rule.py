"""The direct join table's claim rule (one atomic AND per row, a repeated key
ORs back what it saw) under real concurrency on the CPU: tests/native/
claim_stress.c runs it with C11 atomics in 8 threads over a few shared
words, for 1-, 2- and 4-byte entries, and checks that exactly one row per
key claims the entry and that the entry ends holding that row's payload.
The same program without the OR-back must fail, so the check can see the
corruption the OR-back repairs."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def stress(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("claim") / "claim_stress")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-pthread", "-o", exe,
                    os.path.join(HERE, "native", "claim_stress.c")], check=True)
    return exe


@pytest.mark.parametrize("w", [1, 2, 4])
def test_claim_rule_keeps_the_first_rows_payload(stress, w):
    r = subprocess.run([stress, str(w), "400", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "bad_rounds=0" in r.stdout


@pytest.mark.parametrize("w", [1, 4])
def test_claim_rule_without_restore_is_caught(stress, w):
    r = subprocess.run([stress, str(w), "400", "0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 1, r.stdout + r.stderr
