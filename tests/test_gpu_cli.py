"""The command-line programs end to end on the GPU: prerun writes
prerun_results.h5, mainrun (plain loop and --batched) reads it and appends
ln g to mainrun_results.h5 in the reference's layout; the records read back
equal each other and the library path run on the same inputs (couplings,
prerun and mainrun from the same seed, src/prerun.cu and src/mainrun.cu)."""

from __future__ import annotations

import subprocess
import time
from pathlib import Path

import numpy as np
import pytest

from qec_pf_wang_landau_amd import engine

pytestmark = pytest.mark.gpu

BIN = Path(engine.__file__).resolve().parent / "_lib" / "bin"


def run(args, cwd):
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


def dump_results(path: Path):
    recs = []
    for line in run([str(BIN / "results_dump"), str(path)], path.parent).splitlines():
        if line.startswith("#"):
            _, d, s, stamp, n, *ver = line.split()
            recs.append({"disorder": int(d), "seed": int(s), "E": [], "lng": [], "n": int(n), "ver": " ".join(ver)})
        else:
            e, g = line.split()
            recs[-1]["E"].append(int(e))
            recs[-1]["lng"].append(float(g))
    return recs


def test_prerun_mainrun_cli_match_library(cuda, tmp_path):
    if not (BIN / "results_dump").exists():
        pytest.skip("executables built without HDF5")
    common = ["-x", "4", "-y", "4", "-p", "0.1", "-s", "12345", "-e", "v", "-r", "2", "-d", "0"]
    run([str(BIN / "prerun"), *common, "-n", "10000", "-w", "128", "-i", "16"], tmp_path)
    pre = list(tmp_path.rglob("prerun_results.h5"))
    assert len(pre) == 1
    main = ["-n", "1000", "-a", "0.8", "-b", "1e-6", "-i", "4", "-w", "4", "-o", "0.25", "-c", "20", "-f", "3600"]
    run([str(BIN / "mainrun"), *common, *main], tmp_path)
    time.sleep(1.1)  # run groups are named by the second (src/utils.cpp), as in the reference
    run([str(BIN / "mainrun"), *common, *main, "--batched"], tmp_path)
    res = list(tmp_path.rglob("mainrun_results.h5"))
    assert len(res) == 1
    recs = dump_results(res[0])
    # two samples x two runs (plain, batched), in two timestamp groups
    assert len(recs) == 4
    by = {}
    for r in recs:
        assert r["n"] == len(r["E"]) > 0 and r["seed"] == 12345 and r["ver"]
        by.setdefault(r["disorder"], []).append(r)
    for d, rs in by.items():
        assert len(rs) == 2
        assert rs[0]["E"] == rs[1]["E"] and rs[0]["lng"] == rs[1]["lng"]
    # the library path on the same inputs
    cfg = engine.RunConfig(X=4, Y=4, boundary=0, num_intervals=4, walker_per_interval=4, overlap=0.25, alpha=0.8,
                           beta=1e-6, num_iterations=1000, exchange_every=20, seed=12345)
    samples = engine.problem_gpu(cfg, 2, 0.1, "v", prerun_walkers=128, prerun_iters=10000, prerun_intervals=16)
    st = engine.build_state(samples, cfg)
    ds = engine.DeviceState(st)
    rc, _ = engine.run_plain(ds)
    assert rc == 0
    got = ds.pull()
    for s in range(2):
        a, b = got.sample_intervals(s)
        E, lng = engine.stitch(got.sample_logG(s), a, b, cfg.walker_per_interval, cfg.X, cfg.Y)
        assert by[s][0]["E"] == E.tolist()
        assert np.array_equal(np.array(by[s][0]["lng"]), lng)


def prerun_dump(directory: Path, X: int, Y: int) -> dict:
    from test_file_io import parse_prerun_dump  # noqa: PLC0415
    return parse_prerun_dump(run([str(BIN / "prerun_dump"), str(directory), str(X), str(Y)], directory))


def all_energies(J: np.ndarray, X: int, Y: int, boundary: int) -> set:
    """Every energy of the X x Y lattice under couplings J (brute force)."""
    N = X * Y
    codes = np.arange(1 << N, dtype=np.int64)
    spins = (1 - 2 * ((codes[:, None] >> np.arange(N)) & 1)).astype(np.int8)
    return set(engine.lattice_energies(spins, J, X, Y, boundary).tolist())


@pytest.mark.parametrize("boundary", [0, 1])
def test_prerun_files_pass_the_reference_integration_checks(cuda, tmp_path, boundary):
    """The checks of the reference's prerun integration scripts on this
    prerun's HDF5 output: every stored seed lattice has the energy it is
    filed under (test/integration_prerun_check_energies.py), and every
    sample's histogram flags are exactly the energies of all 2^16 lattices
    (test/integration_prerun_matching_hist_energies.py: 4x4, 10 samples,
    10000 steps, 128 walkers, 20 intervals, p in {0, 1})."""
    if not (BIN / "prerun_dump").exists():
        pytest.skip("executables built without HDF5")
    X = Y = 4
    S = 10
    for prob in (0.0, 1.0):
        cwd = tmp_path / f"p{prob}"
        cwd.mkdir()
        run([str(BIN / "prerun"), "-x", str(X), "-y", str(Y), "-p", str(prob), "-n", "10000", "-w", "128", "-s", "99",
             "-i", "20", "-r", str(S), "-d", "99", "--boundary", str(boundary)], cwd)
        files = list(cwd.rglob("prerun_results.h5"))
        assert len(files) == 1
        dump = prerun_dump(files[0].parent, X, Y)
        assert sorted(dump["histogram"]) == list(range(S))
        for s in range(S):
            J = np.array(dump["interaction"][s], np.int8)
            lats = {E: lat for (ss, E), lat in dump["lattice"].items() if ss == s}
            assert lats, f"sample {s}: no seed lattice stored"
            for E, lat in lats.items():
                got = int(engine.lattice_energies(np.array(lat, np.int8)[None, :], J, X, Y, boundary)[0])
                assert got == E, (s, E, got)
            flags = np.array(dump["histogram"][s])
            found = set((np.flatnonzero(flags) - 2 * X * Y).tolist())
            assert found == all_energies(J, X, Y, boundary), s


def test_prerun_coupling_fraction_matches_p(cuda, tmp_path):
    """test/integration_prerun_check_interactions.py: the share of flipped
    couplings of the stored disorder is p within 0.02 (16x16, 20 samples)."""
    if not (BIN / "prerun_dump").exists():
        pytest.skip("executables built without HDF5")
    X = Y = 16
    for boundary in (0, 1):
        for prob in (0.07, 0.31):
            cwd = tmp_path / f"b{boundary}_{prob}"
            cwd.mkdir()
            run([str(BIN / "prerun"), "-x", str(X), "-y", str(Y), "-p", str(prob), "-n", "2000", "-w", "64", "-s",
                 "7", "-i", "16", "-r", "20", "-d", "1", "--boundary", str(boundary)], cwd)
            files = list(cwd.rglob("prerun_results.h5"))
            assert len(files) == 1
            dump = prerun_dump(files[0].parent, X, Y)
            J = np.concatenate([np.array(v) for v in dump["interaction"].values()])
            assert J.size == 20 * 2 * X * Y
            assert abs(np.mean(J == -1) - prob) < 0.02


@pytest.mark.parametrize("boundary", [0, 1])
def test_mainrun_files_pass_the_reference_integration_check(cuda, tmp_path, boundary):
    """test/integration_mainrun.py on this CLI's output (5x5, p = 0.2, 10
    samples; prerun 10000 steps, 128 walkers, 36 intervals; mainrun 1000
    steps, 10 intervals x 4 walkers, alpha 0.8, beta 0.001, overlap 0.25,
    exchange every 20): for every sample, the stored ln g covers exactly the
    prerun's energies and its log-sum-exp is N ln 2."""
    if not (BIN / "prerun_dump").exists():
        pytest.skip("executables built without HDF5")
    X = Y = 5
    S = 10
    common = ["-x", str(X), "-y", str(Y), "-p", "0.2", "-s", "1234", "-e", "h", "-r", str(S), "-d", "99",
              "--boundary", str(boundary)]
    run([str(BIN / "prerun"), *common, "-n", "10000", "-w", "128", "-i", "36"], tmp_path)
    run([str(BIN / "mainrun"), *common, "-n", "1000", "-a", "0.8", "-b", "0.001", "-i", "10", "-w", "4", "-o", "0.25",
         "-c", "20", "-f", "3600", "--batched"], tmp_path)
    pre = list(tmp_path.rglob("prerun_results.h5"))
    res = list(tmp_path.rglob("mainrun_results.h5"))
    assert len(pre) == 1 and len(res) == 1
    dump = prerun_dump(pre[0].parent, X, Y)
    recs = dump_results(res[0])
    assert sorted(r["disorder"] for r in recs) == list(range(S))
    for r in recs:
        flags = np.array(dump["histogram"][r["disorder"]])
        assert set(r["E"]) == set((np.flatnonzero(flags) - 2 * X * Y).tolist())
        g = np.array(r["lng"])
        m = g.max()
        assert m + np.log(np.exp(g - m).sum()) == pytest.approx(X * Y * np.log(2.0), abs=1e-10)
