"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The plain-C host sources -- the MPS reader (scip_amd/csrc/mps.c), the generators
(scip_amd/csrc/gen.c) and the oracle (oracle/oracle.c) -- are compiled with gcc
-fsanitize=address,undefined together with tests/sanitize_driver.c, which reads the 12 MIPLIB
fixtures and a set of malformed files, generates instances with every generator and propagates
them with both oracle schedules (with and without ranged rows and tightening).  Any sanitizer
report fails the run.  The HIP engine and the C++ server are device code or drive it and are
not part of this pass.
"""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MPS = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "mps", "*.mps")))

MALFORMED = {
    "bad_rowtype.mps": "NAME X\nROWS\n Q  r1\nENDATA\n",
    "no_endata.mps": "NAME X\nROWS\n N  obj\n L  r1\nCOLUMNS\n    x  r1  1.0\n",
    "unknown_row.mps": "NAME X\nROWS\n N  obj\n L  r1\nCOLUMNS\n    x  r9  1.0\nRHS\nENDATA\n",
    "bad_number.mps": "NAME X\nROWS\n N  obj\n L  r1\nCOLUMNS\n    x  r1  1.0zz\nENDATA\n",
    "empty.mps": "",
    "long_line.mps": "NAME " + "X" * 5000 + "\nROWS\n N  obj\n L  " + "r" * 3000 + "\nENDATA\n",
    "bounds_unknown_col.mps": "NAME X\nROWS\n N  obj\n L  r1\nCOLUMNS\n    x  r1  1.0\nRHS\n    RHS  r1  2\n"
                              "BOUNDS\n UP BND  y  3\nENDATA\n",
    "ranges_only.mps": "NAME X\nROWS\n N  obj\n E  r1\nCOLUMNS\n    x  r1  1.0\nRANGES\n    RNG  r1  -4\n"
                       "ENDATA\n",
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_host_sources_clean_under_asan_and_ubsan(tmp_path):
    exe = tmp_path / "sanitize_driver"
    cmd = ["gcc", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(ROOT, 'oracle')}", os.path.join(ROOT, "tests", "sanitize_driver.c"),
           os.path.join(ROOT, "scip_amd", "csrc", "mps.c"), os.path.join(ROOT, "scip_amd", "csrc", "gen.c"),
           os.path.join(ROOT, "oracle", "oracle.c"), "-lm", "-o", str(exe)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    files = list(MPS)
    for name, text in MALFORMED.items():
        f = tmp_path / name
        f.write_text(text)
        files.append(str(f))
    files.append(str(tmp_path / "missing.mps"))
    # verify_asan_link_order=0: the environment may preload a library of its own
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")
    r = subprocess.run([str(exe)] + files, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "sanitize driver done" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout
    # the fixtures read; broken structure is refused, while entries for unknown rows or
    # columns are skipped and numbers are read as far as they parse, as SCIP's reader does
    # (reader_mps.c:1159-1169, :2084, atof)
    for p in MPS:
        assert f"{p}: rc 1 " in r.stdout
    refused = {"bad_rowtype.mps", "no_endata.mps", "empty.mps"}
    for name in MALFORMED:
        line = next(l for l in r.stdout.splitlines() if l.startswith(str(tmp_path / name) + ":"))
        assert (" rc 1 " not in line) == (name in refused), line
    assert f"{tmp_path / 'missing.mps'}: rc -4 " in r.stdout
