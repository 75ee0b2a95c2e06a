"""The dword split of the wave-cooperative particle-row transfers, checked on
the host by compiling tests/rowspan_check.cpp against the kernels' header."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_row_span_covers_rows_exactly_once():
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "rowspan_check")
        subprocess.run([cxx, "-std=c++17", "-O1", os.path.join(HERE, "rowspan_check.cpp"), "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
