"""The oracle under AddressSanitizer and UBSan: `make -C oracle sanitize`
builds the oracle with oracle/sanitize_check.cpp, which runs every entry
point on small scenes with edge inputs (culled and off-screen Gaussians,
inverted and NaN boxes, empty tiles, threaded adjoint), and aborts on any
out-of-bounds access, use after free, leak or undefined operation."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


def _asan_available():
    cxx = shutil.which(os.environ.get("CXX", "g++"))
    if not cxx:
        return False
    r = subprocess.run([cxx, "-fsanitize=address,undefined", "-x", "c++", "-", "-o", os.devnull],
                       input="int main(){return 0;}", capture_output=True, text=True)
    return r.returncode == 0


@pytest.mark.skipif(not _asan_available(), reason="no AddressSanitizer runtime for the host compiler")
def test_oracle_clean_under_sanitizers():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True,
                       timeout=600)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "sanitize_check: ok" in r.stdout
