"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
edge-list loader (well-formed files and the malformed inputs it must reject:
ids >= V, negative ids, unparsable numbers), the CSR build, the generators,
the walk-query extractor, the two-hop index builder, the score order and the
CPU oracle / trie baseline, all driven by tests/sanitize/driver.cpp."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_and_ubsan(tmp_path):
    out = tmp_path / "build"
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize"), f"OUT={out}"], check=True,
                   capture_output=True, timeout=600)
    files = []
    cases = {
        "plain.g": ("4\n0\t1\n1\t2\n2\t3\n3\t0\n", "ok"),
        "dups_loops_junk.g": ("5\n0\t1\n0\t1\n2\t2\nnot an arc\n  3\t 4\n1\t0\n\n", "ok"),
        "no_arcs.g": ("3\n", "ok"),
        "id_too_large.g": ("3\n0\t3\n", "err"),
        "negative.g": ("3\n-1\t2\n", "err"),
        "garbage_number.g": ("3\nx\t1\n", "err"),
        "empty.g": ("", "err"),
    }
    for name, (text, want) in cases.items():
        p = tmp_path / name
        p.write_text(text)
        files += [str(p), want]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=3",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(out / "driver")] + files, capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitize driver ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
