"""The hand-worked pins of the oracle's texture combiner, dust_remove, debayer, fat zero and
centre of mass (tests/test_oracle.py) fail on plausible mistakes.  Each case builds a copy of the
oracle with one mistake planted (a scratch directory, CPU only) and runs those pins against it;
at least one must fail."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PINS = "combine_rgba or dust_remove or debayer_delta or fat_zero or max_x_cm"

MUTANTS = {
    "final mix of the debayered tiles": ("s += (real)(sw * pw[ip][i]) * iclt[ip][chn][i];",
                                         "s += (real)(sw * pw[ip][i]) * deb[ip][chn][i];"),
    "ksigma not halved": ("real ksigma = (real)(1.0 / (2.0 * prm->diff_sigma * prm->diff_sigma));",
                          "real ksigma = (real)(1.0 / (prm->diff_sigma * prm->diff_sigma));"),
    "dust_remove average from port 0": ("            if (ip != wi) avg += pw[ip][i];",
                                        "            if (ip != wi) avg += pw[ip][i] + ((ip == 1 && wi != 0) ? pw[0][i] : 0);"),
    "second best port: last of a tie": ("if ((ip != b1) && (pw[ip][i] > pw[b2][i])) b2 = ip;",
                                        "if ((ip != b1) && (pw[ip][i] >= pw[b2][i])) b2 = ip;"),
    "fat zero not squared": ("real fz2 = sa2 / 64 * fat_zero * fat_zero;", "real fz2 = sa2 / 64 * fat_zero;"),
    "green side factor 7/6": ("real k_side = rb ? (real)(4.0 / 3.0) : (real)(8.0 / 7.0);",
                              "real k_side = rb ? (real)(4.0 / 3.0) : (real)(7.0 / 6.0);"),
    "centre of mass keeps negative samples": ("if (!isnan(d) && (d > 0)) {", "if (!isnan(d)) {"),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_pins_catch_mutant(tmp_path, name):
    old, new = MUTANTS[name]
    for sub in ("oracle", "include", "tests"):
        shutil.copytree(os.path.join(ROOT, sub), tmp_path / sub,
                        ignore=shutil.ignore_patterns("*.so", "__pycache__", "test_oracle_mutants.py"))
    src_path = tmp_path / "oracle" / "clt_oracle.c"
    src = src_path.read_text()
    assert src.count(old) == 1, "the mutated line is not in the oracle: " + name
    src_path.write_text(src.replace(old, new))
    subprocess.run(["make", "-s", "-C", str(tmp_path / "oracle"), "libclt_oracle64.so"], check=True)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-k", PINS,
                        str(tmp_path / "tests" / "test_oracle.py")], capture_output=True, text=True, cwd=str(tmp_path))
    assert r.returncode == 1 and " failed" in r.stdout, r.stdout[-2000:]
