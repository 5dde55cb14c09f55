"""The library's host code (Matrix Market loader and writer, COO sort and
containers, generators, the 32-bit limit check) and the CPU oracle under
AddressSanitizer + LeakSanitizer and UBSan, built with g++/gcc on the CPU
(no HIP code is compiled here).  tests/sanitize_driver.cpp reads every
golden .mtx file and a set of malformed ones, round-trips them through
spgWriteMTX, and runs every oracle phase; any sanitizer report aborts it."""
import glob
import os
import shutil
import subprocess

import pytest

from helpers import GOLDEN

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g",
       "-O1"]

MALFORMED = {
    "empty.mtx": "",
    "banner_only.mtx": "%%MatrixMarket matrix coordinate real general\n",
    "bad_banner.mtx": "%%MatrixMarket tensor coordinate real general\n2 2 1\n1 1 1.0\n",
    "truncated.mtx": "%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1.0\n2 2 2.0\n",
    "out_of_range.mtx": "%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1.0\n",
    "zero_index.mtx": "%%MatrixMarket matrix coordinate real general\n3 3 1\n0 1 1.0\n",
    "negative_size.mtx": "%%MatrixMarket matrix coordinate real general\n-3 3 1\n1 1 1.0\n",
    "garbage_value.mtx": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 abc\n",
    "huge_line.mtx": "%%MatrixMarket matrix coordinate real general\n%" + "x" * 100000
                     + "\n2 2 1\n1 2 3.5\n",
    "symmetric_upper.mtx": "%%MatrixMarket matrix coordinate real symmetric\n3 3 2\n1 2 1.0\n3 3 2.0\n",
}


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None,
                    reason="host compilers absent")
def test_host_code_and_oracle_under_asan_ubsan(tmp_path):
    obj = str(tmp_path / "oracle.o")
    exe = str(tmp_path / "sanitize_driver")
    subprocess.run(["gcc", "-std=gnu11", "-c", os.path.join(ROOT, "oracle", "spgemm_oracle.c"),
                    "-o", obj] + SAN, check=True)
    subprocess.run(["g++", "-std=c++17", os.path.join(ROOT, "tests", "sanitize_driver.cpp"),
                    os.path.join(ROOT, "spgemm_amd", "csrc", "host_formats.cpp"), obj, "-o", exe,
                    "-pthread", "-lm"] + SAN, check=True)
    bad = []
    for name, text in MALFORMED.items():
        p = tmp_path / name
        p.write_text(text)
        bad.append(str(p))
    files = sorted(glob.glob(os.path.join(GOLDEN, "*.mtx"))) + bad
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe, str(tmp_path)] + files, capture_output=True, text=True, env=env,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "0 failed checks" in r.stdout
