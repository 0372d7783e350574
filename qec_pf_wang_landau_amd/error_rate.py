"""Logical error rates of the surface code under independent bit flips,
decoded by Wang-Landau partition functions: the Monte-Carlo step after the
class decision (README.md:198-200 cites the paper whose end product these
curves are; the reference stops at per-class ln g files).

For one point (L, p): S disorder samples (the couplings of sample q come from
seed + q, src/wl.cu:380), each with its four error classes i/v/h/c
(src/prerun.cu:52-75).  All 4S class tasks run on the batched path to
ln f <= beta, are stitched and normalised on the device
(src/result_handling.cpp:368-431) and give ln Z_c at the Nishimori
temperature.  Class i is the coset of the sampled error itself, so a sample
fails unless i is the most likely class; the rate is the mean failure over
the decoded samples, with a Wilson score interval.

``method="exact"`` replaces the Wang-Landau estimate by the exact density of
states (Gray-code enumeration on the GPU, N <= 40): the same samples, the
exact decisions.  ``decode_result_files`` computes the same point from four
mainrun_results.h5 files (this library's CLI or the reference's).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import subprocess
import sys
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import engine
from ._lib import LIB_PATH, call

# Lattice size -> (intervals, walkers per interval): every class task of these
# sizes fits one workgroup's LDS on the batched kernel.
DEFAULT_SIZES = {4: (4, 8), 5: (6, 8), 6: (10, 8), 8: (12, 8), 10: (14, 8), 12: (16, 8), 16: (32, 8)}

DECODE_RESULTS = LIB_PATH.parent / "bin" / "decode_results"


@dataclass
class Point:
    """Decoding statistics of one (L, p) point."""
    L: int
    p: float
    method: str
    samples: int            # samples requested
    decoded: int            # samples whose four classes all finished
    failures: float         # sum of per-sample failure weights
    rate: float
    lo: float               # Wilson interval at z
    hi: float
    z: float
    log_z: np.ndarray       # [S, 4] ln Z per class i/v/h/c (NaN when undecoded)
    decision: np.ndarray    # [S] argmax class, -1 when undecoded
    status: np.ndarray      # [S, 4] task status (1 = finished)
    seconds: float
    attempts: int = 0
    extra: dict = field(default_factory=dict)

    def summary(self) -> dict:
        return {"L": self.L, "p": self.p, "method": self.method, "samples": self.samples,
                "decoded": self.decoded, "failures": self.failures, "rate": self.rate, "wilson": [self.lo, self.hi],
                "z": self.z, "seconds": self.seconds, "attempts": self.attempts, **self.extra}


def decide(log_z: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """wl_decide_classes: per sample the first argmax class (-1 when a ln Z is
    not finite) and the failure weight (1 - 1/m among m tied maxima that
    include class i)."""
    z = np.ascontiguousarray(log_z, np.float64)
    S, K = z.shape
    dec = np.zeros(S, np.int32)
    fail = np.zeros(S, np.float64)
    call("wl_decide_classes", z.ctypes.data_as(C.c_void_p), S, K, dec.ctypes.data_as(C.c_void_p),
         fail.ctypes.data_as(C.c_void_p))
    return dec, fail


def wilson(failures: float, n: int, z: float = 1.96) -> tuple[float, float]:
    """wl_wilson_interval: the Wilson score interval of failures / n."""
    lo, hi = C.c_double(), C.c_double()
    call("wl_wilson_interval", float(failures), int(n), float(z), C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def _aggregate(L, p, method, S, log_z, status, seconds, z, attempts=0, extra=None) -> Point:
    ok = (status == 1).all(axis=1)
    lz = np.where(ok[:, None], log_z, np.nan)
    dec, fail = decide(lz)
    decided = dec >= 0
    n = int(decided.sum())
    failures = float(fail[decided].sum())
    lo, hi = wilson(failures, n, z) if n else (0.0, 1.0)
    return Point(L, p, method, S, n, failures, failures / n if n else float("nan"), lo, hi, z, lz, dec, status,
                 seconds, attempts, extra or {})


def run_config(L: int, boundary: int = 0, seed: int = 12345, num_intervals: int | None = None,
               walker_per_interval: int | None = None, beta: float = 1e-6, alpha: float = 0.8,
               num_iterations: int = 1000, exchange_every: int = 20, overlap: float = 0.25) -> engine.RunConfig:
    I, W = DEFAULT_SIZES.get(L, (max(4, L + L // 2), 8))
    return engine.RunConfig(X=L, Y=L, boundary=boundary, num_intervals=num_intervals or I,
                            walker_per_interval=walker_per_interval or W, overlap=overlap, alpha=alpha, beta=beta,
                            num_iterations=num_iterations, exchange_every=exchange_every, seed=seed)


def decode_point(L: int, p: float, samples: int, cfg: engine.RunConfig | None = None, method: str = "wl",
                 sample_base: int = 0, prerun_walkers: int = 128, prerun_iters: int | None = None,
                 max_rounds: int = 1 << 40, z: float = 1.96) -> Point:
    """Decode `samples` disorder samples of an L x L lattice at flip
    probability p and aggregate the logical error rate.  Sample q uses
    couplings and walker streams of cfg.seed + sample_base + q."""
    torch = engine._torch()
    cfg = cfg or run_config(L)
    if cfg.X != L or cfg.Y != L:
        raise ValueError("cfg must describe an L x L lattice")
    S = int(samples)
    if S < 1:
        raise ValueError("need at least one sample")
    bN = engine.nishimori_beta(p)
    t0 = time.perf_counter()
    if method == "exact":
        N = L * L
        if N > 40:
            raise ValueError("exact decoding enumerates 2^N states: N = L^2 <= 40")
        seed = cfg.seed + sample_base
        J = np.concatenate([engine.disorder_gpu(seed, S, L, L, p, c) for c in engine.ERROR_CLASSES])
        counts = engine.exact_counts_gpu(J, L, L, cfg.boundary)
        torch.cuda.synchronize()
        energies = np.arange(-2 * N, 2 * N + 1, dtype=np.int32)
        lz = np.zeros(4 * S)
        for q in range(4 * S):
            nz = counts[q] != 0
            lz[q] = engine.log_partition(energies[nz], np.log(counts[q][nz].astype(np.float64)), bN)
        log_z = lz.reshape(4, S).T
        return _aggregate(L, p, method, S, log_z, np.ones((S, 4), np.int32), time.perf_counter() - t0, z)
    if method != "wl":
        raise ValueError("method must be 'wl' or 'exact'")
    states = class_states(cfg, p, S, sample_base, prerun_walkers, prerun_iters)
    return decode_states(states, p, max_rounds, z, t0)


def class_states(cfg: engine.RunConfig, p: float, samples: int, sample_base: int = 0, prerun_walkers: int = 128,
                 prerun_iters: int | None = None) -> list[engine.MainrunState]:
    """Couplings and prerun on the GPU, then the mainrun state of each error
    class i, v, h, c over the same samples (src/prerun.cu, src/mainrun.cu:42-228)."""
    iters = prerun_iters or 2000 * cfg.X * cfg.Y
    states = []
    for c in engine.ERROR_CLASSES:
        inputs = engine.problem_gpu(cfg, samples, p, c, prerun_walkers=prerun_walkers, prerun_iters=iters,
                                    prerun_intervals=max(16, 2 * cfg.num_intervals), sample_base=sample_base)
        states.append(engine.build_state(inputs, cfg))
    return states


def decode_states(states: list[engine.MainrunState], p: float, max_rounds: int = 1 << 40, z: float = 1.96,
                  t0: float | None = None) -> Point:
    """Runs the four class states (i, v, h, c; S samples each) as one batch
    to convergence, stitches and takes ln Z at the Nishimori temperature of p
    on the device, and aggregates the rate."""
    torch = engine._torch()
    t0 = time.perf_counter() if t0 is None else t0
    if len(states) != 4 or len({st.S for st in states}) != 1:
        raise ValueError("need the states of the four classes over the same samples")
    S = states[0].S
    L = states[0].cfg.X
    bN = engine.nishimori_beta(p)
    batch = engine.merge_states(states)
    dev = engine.BatchDevice(batch)
    form = "fused" if all(engine.state_fits_one_workgroup(st) for st in states) else "block"
    if form == "fused":
        dev.run(max_rounds)
    else:
        dev.run_blocks(max_rounds)
    dec = engine.stitch_tasks_gpu(dev, np.full(batch.tasks.size, bN))
    res = dev.pull()
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    status = res.arrays["status"].reshape(4, S).T
    rounds = res.arrays["rounds"].astype(np.int64)
    attempts = int(np.sum(rounds * (batch.tasks["I"] * batch.tasks["W"]).astype(np.int64)
                          * batch.tasks["num_iterations"].astype(np.int64)))
    return _aggregate(L, p, "wl", S, dec.log_z.reshape(4, S).T, status, seconds, z, attempts,
                      {"form": form, "rounds_max": int(rounds.max()), "rounds_mean": float(rounds.mean())})


def error_curve(Ls, ps, samples: int, method: str = "wl", **kw) -> list[Point]:
    """decode_point over every (L, p) of the grid, L-major."""
    return [decode_point(L, p, samples, method=method, **kw) for L in Ls for p in ps]


def crossing(a: list[Point], b: list[Point]) -> float | None:
    """p where the rate curves of two sizes cross (linear interpolation
    between the first pair of neighbouring p whose rate difference changes
    sign), or None.  Both lists must share the same ascending p grid."""
    pa = [pt.p for pt in a]
    if pa != [pt.p for pt in b]:
        raise ValueError("curves must share the p grid")
    d = [pa_.rate - pb_.rate for pa_, pb_ in zip(a, b)]
    for k in range(len(d) - 1):
        if d[k] == 0.0:
            return pa[k]
        if d[k] * d[k + 1] < 0.0:
            return pa[k] + (pa[k + 1] - pa[k]) * d[k] / (d[k] - d[k + 1])
    return pa[-1] if d and d[-1] == 0.0 else None


def decode_result_files(p: float, files: list[str | Path], z: float = 1.96, binary: Path = DECODE_RESULTS) -> dict:
    """The decode_results executable over the four mainrun_results.h5 files of
    classes i, v, h, c: per-disorder ln Z and decision, and the rate."""
    if len(files) != 4:
        raise ValueError("need the results files of the four classes i, v, h, c")
    if not binary.exists():
        raise FileNotFoundError(f"{binary} is missing: build the package first")
    out = subprocess.run([str(binary), "--prob", repr(float(p)), "--z", repr(float(z)), *map(str, files)],
                         capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"decode_results failed: {out.stderr.strip()}")
    ids, lz, dec = [], [], []
    summary = {}
    for line in out.stdout.splitlines():
        parts = line.split()
        if line.startswith("#"):
            it = iter(parts[1:])
            for k in it:
                v = next(it)
                summary[k] = [float(v), float(next(it))] if k == "wilson" else float(v)
            continue
        ids.append(int(parts[0]))
        lz.append([float(x) for x in parts[1:5]])
        dec.append("ivhc".index(parts[5]) if parts[5] != "-" else -1)
    return {"disorder": np.array(ids, np.int32), "log_z": np.array(lz, np.float64).reshape(-1, 4),
            "decision": np.array(dec, np.int32), **summary}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Logical error rates from Wang-Landau decoding on the GPU")
    ap.add_argument("--L", type=int, nargs="+", default=[4, 6], help="lattice sizes (L x L)")
    ap.add_argument("--p", type=float, nargs="+", default=[0.04, 0.08, 0.12], help="bit-flip probabilities")
    ap.add_argument("--samples", type=int, default=100, help="disorder samples per (L, p) point")
    ap.add_argument("--seed", type=int, default=12345, help="seed of sample 0 (sample q uses seed + q)")
    ap.add_argument("--boundary", type=int, choices=(0, 1), default=0, help="0 periodic, 1 open")
    ap.add_argument("--beta", type=float, default=1e-6, help="final ln f")
    ap.add_argument("--method", choices=("wl", "exact"), default="wl",
                    help="Wang-Landau or exact enumeration (N <= 40)")
    args = ap.parse_args(argv)
    for L in args.L:
        cfg = run_config(L, boundary=args.boundary, seed=args.seed, beta=args.beta)
        for p in args.p:
            pt = decode_point(L, p, args.samples, cfg, method=args.method)
            print(json.dumps(pt.summary()), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
