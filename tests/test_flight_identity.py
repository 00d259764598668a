"""The step kernel's fused free-flight draw equals the reference's
1 - xi for every LCG state (tests/cpp/flight_identity.c, all 2^32 states,
about 15 s on one core)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_fused_one_minus_xi_over_all_states(tmp_path):
    exe = os.path.join(str(tmp_path), "flight_identity")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "flight_identity.c"), "-o", exe, "-lm"], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert r.returncode == 0, r.stdout
    assert "0 of 4294967296 states differ" in r.stdout
