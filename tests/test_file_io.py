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


def parse_prerun_dump(text: str) -> dict:
    """Sections of prerun_dump's output: histogram / interaction per sample,
    lattice per (sample, energy)."""
    out = {"histogram": {}, "interaction": {}, "lattice": {}}
    lines = text.splitlines()
    for head, body in zip(lines[::2], lines[1::2]):
        _, kind, s, x = head.split()
        if kind == "histogram":
            out[kind][int(s)] = [int(c) for c in body]
        elif kind == "interaction":
            out[kind][int(s)] = [int(v) for v in body.split()]
        else:
            out[kind][(int(s), int(x))] = [int(v) for v in body.split()]
    return out


def test_prerun_dump_reads_the_reference_vectors(tmp_path):
    """prerun_dump prints the prerun file io_check wrote from the reference's
    file-test vectors (test/test_file_io.cpp)."""
    test_file_layer_matches_reference_vectors(tmp_path)
    exe = build.OUT / "bin" / "prerun_dump"
    if not exe.exists():
        pytest.skip("executables built without HDF5")
    run = subprocess.run([str(exe), str(tmp_path / "scratch" / "prerun"), "4", "4", "5"], capture_output=True,
                         text=True, timeout=60)
    assert run.returncode == 0, run.stderr
    d = parse_prerun_dump(run.stdout)
    assert d["histogram"] == {0: [1, 0, 0, 0, 1], 1: [0, 0, 1, 0, 1]}
    assert d["interaction"][0][:4] == [1, -1, 1, 1] and len(d["interaction"][1]) == 32
    assert sorted(d["lattice"]) == [(0, 0), (0, 10), (0, 20), (1, 40), (1, 50), (1, 70)]
    assert d["lattice"][(0, 10)] == [1, 1, 1, -1, 1, 1, 1, -1, -1, 1, 1, 1, 1, 1, 1, 1]
    bad = subprocess.run([str(exe), str(tmp_path), "4", "4"], capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0
