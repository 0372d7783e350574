"""The prerun/mainrun file layer (csrc/wl_io.cpp) against the reference's
file-test vectors and path scheme: compiles tests/cpp/io_check.cpp with the
file layer and HDF5 on the host and runs it (no GPU)."""

from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

from qec_pf_wang_landau_amd import build

ROOT = Path(__file__).resolve().parents[1]


def test_file_layer_matches_reference_vectors(tmp_path):
    flags = build.hdf5_flags()
    if flags is None:
        pytest.skip("HDF5 is not available on this host")
    cflags, _ = flags
    priv = build.OUT / "hdf5"  # the private directory hdf5_flags links through
    exe = tmp_path / "io_check"
    cmd = ["g++", "-std=c++17", "-O1", f"-I{build.CSRC}", *cflags, str(ROOT / "tests" / "cpp" / "io_check.cpp"),
           str(build.CSRC / "wl_io.cpp"), "-o", str(exe), f"-L{priv}", "-lhdf5", f"-Wl,-rpath,{priv}"]
    comp = subprocess.run(cmd, capture_output=True, text=True)
    assert comp.returncode == 0, comp.stderr
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    run = subprocess.run([str(exe), str(scratch)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "all checks passed" in run.stdout
