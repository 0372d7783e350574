"""GPU checks of the logical-error-rate pipeline (error_rate.py):

* the Wang-Landau point (four classes as one batch, device stitching and
  ln Z) equals, bit for bit, the oracle running every class task alone,
  stitching and taking ln Z on the CPU; so do the decisions and the rate;
* the exact point (Gray-code DOS on the GPU) equals the oracle's
  transfer-matrix counts taken through the oracle's ln Z;
* Wang-Landau decisions equal the exact decisions wherever the exact ln Z
  gap between the two most likely classes exceeds 0.5, and the two rates
  differ by at most the share of samples without such a gap;
* the command-line path (prerun, mainrun per class, decode_results on the
  four result files) gives the library path's ln Z and decisions."""

from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from helpers import oracle_per_task, oracle_samples
from oracle_lib import exact_lng, oracle, p
from qec_pf_wang_landau_amd import engine, error_rate

pytestmark = pytest.mark.gpu

CLASSES = engine.ERROR_CLASSES
BIN = Path(engine.__file__).resolve().parent / "_lib" / "bin"


def oracle_log_z(st: engine.MainrunState, beta: float) -> np.ndarray:
    o = oracle()
    cfg = st.cfg
    out = np.zeros(st.S)
    for s, ref in enumerate(oracle_per_task(st, 1 << 40)):
        assert ref.cond_sample[0] == -1, f"oracle task {s} did not finish"
        a, b = ref.sample_intervals(0)
        cap = int(b[-1] - a[0] + 1)
        oE = np.zeros(cap, np.int32)
        og = np.zeros(cap)
        n = o.wlo_stitch(p(np.ascontiguousarray(ref.sample_logG(0))), p(np.ascontiguousarray(a)),
                         p(np.ascontiguousarray(b)), cfg.num_intervals, cfg.walker_per_interval, cfg.X, cfg.Y,
                         p(oE), p(og), cap)
        assert n > 0
        out[s] = o.wlo_log_partition(p(oE), p(og), n, beta)
    return out


def expected_failures(lz: np.ndarray) -> float:
    total = 0.0
    for row in lz:
        top = np.flatnonzero(row == row.max())
        total += 1.0 - (1.0 / top.size if 0 in top else 0.0)
    return total


@pytest.mark.parametrize("L,boundary,prob", [(4, 0, 0.1), (5, 1, 0.08)])
def test_wl_point_equals_oracle(cuda, L, boundary, prob):
    S = 6
    cfg = error_rate.run_config(L, boundary=boundary, seed=777)
    states = [engine.build_state(oracle_samples(cfg, S, prob, c), cfg) for c in CLASSES]
    pt = error_rate.decode_states([st.copy() for st in states], prob)
    assert pt.decoded == S and (pt.status == 1).all()
    bN = engine.nishimori_beta(prob)
    ref = np.stack([oracle_log_z(st, bN) for st in states], axis=1)
    assert np.array_equal(pt.log_z, ref)
    assert pt.decision.tolist() == np.argmax(ref, axis=1).tolist()
    assert pt.failures == expected_failures(ref)
    assert pt.rate == pt.failures / S
    assert pt.lo <= pt.rate <= pt.hi


def test_exact_point_equals_oracle_transfer_matrix(cuda):
    L, S, prob = 4, 10, 0.12
    cfg = error_rate.run_config(L, seed=4242)
    pt = error_rate.decode_point(L, prob, S, cfg, method="exact")
    assert pt.decoded == S
    o = oracle()
    bN = engine.nishimori_beta(prob)
    N = L * L
    ref = np.zeros((S, 4))
    for ci, c in enumerate(CLASSES):
        J = engine.disorder_gpu(cfg.seed, S, L, L, prob, c)
        for s in range(S):
            lng = exact_lng(J[s * 2 * N:(s + 1) * 2 * N], L, L, 0)
            fin = np.isfinite(lng)
            E = (np.flatnonzero(fin) - 2 * N).astype(np.int32)
            g = np.ascontiguousarray(lng[fin])
            ref[s, ci] = o.wlo_log_partition(p(E), p(g), E.size, bN)
    assert np.allclose(pt.log_z, ref, rtol=0, atol=1e-12)
    assert pt.decision.tolist() == np.argmax(ref, axis=1).tolist()


def test_wl_decisions_match_exact_decisions(cuda):
    """6x6 periodic at p = 0.1 (near the threshold, so failures occur):
    24 samples, Wang-Landau to ln f = 1e-6 against exact enumeration."""
    L, S, prob = 6, 24, 0.1
    cfg = error_rate.run_config(L, seed=2025)
    wl = error_rate.decode_point(L, prob, S, cfg)
    ex = error_rate.decode_point(L, prob, S, cfg, method="exact")
    assert wl.decoded == S and ex.decoded == S
    srt = np.sort(ex.log_z, axis=1)
    clear = (srt[:, -1] - srt[:, -2]) > 0.5
    assert clear.sum() >= S // 2
    assert (wl.decision[clear] == ex.decision[clear]).all()
    assert abs(wl.rate - ex.rate) <= (S - clear.sum()) / S + 1e-12
    assert ex.failures > 0  # the point is informative


def test_cli_results_files_decode_like_the_library(cuda, tmp_path):
    exe = BIN / "decode_results"
    if not exe.exists():
        pytest.skip("executables built without HDF5")
    prob, S = 0.1, 3
    common = ["-x", "4", "-y", "4", "-p", str(prob), "-s", "321", "-r", str(S), "-d", "0"]
    main = ["-n", "1000", "-a", "0.8", "-b", "1e-6", "-i", "4", "-w", "8", "-o", "0.25", "-c", "20", "-f", "3600",
            "--batched"]
    files = []
    for c in CLASSES:
        for prog, extra in (("prerun", ["-n", "10000", "-w", "128", "-i", "16"]), ("mainrun", main)):
            r = subprocess.run([str(BIN / prog), *common, "-e", c, *extra], cwd=tmp_path, capture_output=True,
                               text=True, timeout=240)
            assert r.returncode == 0, r.stdout + r.stderr
        found = [f for f in tmp_path.rglob("mainrun_results.h5") if f.parent.name == f"error_class_{c}"]
        assert len(found) == 1
        files.append(found[0])
    out = error_rate.decode_result_files(prob, files)
    assert out["disorder"].tolist() == list(range(S)) and out["samples"] == S and out["missing"] == 0
    # the library path on the same inputs
    cfg = engine.RunConfig(X=4, Y=4, boundary=0, num_intervals=4, walker_per_interval=8, overlap=0.25, alpha=0.8,
                           beta=1e-6, num_iterations=1000, exchange_every=20, seed=321)
    states = [engine.build_state(engine.problem_gpu(cfg, S, prob, c, prerun_walkers=128, prerun_iters=10000,
                                                    prerun_intervals=16), cfg) for c in CLASSES]
    pt = error_rate.decode_states(states, prob)
    assert np.array_equal(out["log_z"], pt.log_z)
    assert out["decision"].tolist() == pt.decision.tolist()
    assert out["failures"] == pt.failures and out["rate"] == pt.rate
