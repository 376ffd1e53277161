"""Mutation check of the oracle's string-iteration and FreeE pins.

Each mutant rebuilds the oracle (oracle/scft_oracle.c) with ORC_MUTANT = k, which plants
one plausible mistake, and runs tests/test_oracle_string_pins.py against that build in a
fresh process (ORC_LIB). The pins must fail for every mutant; the unmutated oracle passes
them in the ordinary run of the suite.

Not listed: redistribution applied to the end replicas as well. The natural spline
interpolates its nodes and x[0] = xref[0] = 0, x[R-1] = xref[R-1] = 1, so that update is
(a_end - a_end) * lambda, zero up to rounding: the mutant is equivalent to the
restatement and no test can tell them apart.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "scft_oracle.c")
PINS = os.path.join(ROOT, "tests", "test_oracle_string_pins.py")

MUTANTS = {
    1: "mixing with lambda 0.5 instead of 0.05 (scft.cu:562, 690-693)",
    2: "errs and errs3 normalised by M instead of 2M (scft.cu:630, 674)",
    3: "arc length over W- and W+ instead of W- only (scft.cu:351-355)",
    4: "redistribution skipped (scft.cu:745-757)",
    6: "spring constant 1.0 instead of 0.5 (scft.cu:570, 606)",
    7: "projection adds the tangential part instead of removing it (scft.cu:672)",
    8: "redistribution with lambda 0.5 (density.h:527)",
    9: "FreeE with +wA phiA instead of -wA phiA (scft.cu:825)",
}


def _run_mutant(k, tmp):
    so = os.path.join(tmp, f"liborc_mutant{k}.so")
    b = subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=c99", f"-DORC_MUTANT={k}", "-shared", "-o", so,
                        SRC, "-lm"], capture_output=True, text=True)
    if b.returncode != 0:
        return k, None, b.stderr
    env = dict(os.environ, ORC_LIB=so, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", PINS, "-q", "-p", "no:cacheprovider"], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=300)
    return k, r.returncode, r.stdout[-3000:]


def test_every_mutant_fails_the_pins(tmp_path):
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(lambda k: _run_mutant(k, str(tmp_path)), sorted(MUTANTS)))
    survived = []
    for k, rc, out in results:
        assert rc is not None, f"mutant {k} did not build:\n{out}"
        failed = [l.split("::")[1].split(" ")[0] for l in out.splitlines() if l.startswith("FAILED")]
        print(f"mutant {k} ({MUTANTS[k]}): rc={rc}, failing pins: {failed}")
        if rc != 1 or not failed:   # 1 = tests ran and some failed
            survived.append((k, MUTANTS[k], rc, out[-800:]))
    assert not survived, survived


@pytest.mark.parametrize("k", [0])
def test_unmutated_build_passes(tmp_path, k):
    """ORC_MUTANT = 0 compiles to the restatement itself."""
    _, rc, out = _run_mutant(k, str(tmp_path))
    assert rc == 0, out
