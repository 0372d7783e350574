"""Builders of test problems: disorder, prerun and mainrun state, driven by the
CPU oracle so that CPU-only tests and GPU parity tests start from the same
inputs."""

from __future__ import annotations

import numpy as np

from oracle_lib import oracle, p
from qec_pf_wang_landau_amd import engine


def oracle_disorder(seed: int, S: int, X: int, Y: int, prob: float, cls: str) -> np.ndarray:
    """src/prerun.cu:52-75: couplings of S samples with the class chain applied."""
    N = X * Y
    J = np.zeros(S * 2 * N, np.int8)
    o = oracle()
    o.wlo_init_interactions(p(J), N, S, seed, prob)
    if cls in ("h", "c"):
        o.wlo_apply_x_horizontal_error(p(J), X, Y, S)
    if cls in ("v", "c"):
        o.wlo_apply_x_vertical_error(p(J), X, Y, S)
    return J


def oracle_prerun(J: np.ndarray, X: int, Y: int, boundary: int, seed: int, S: int, walkers: int, iters: int,
                  num_intervals: int):
    """src/prerun.cu:25-133 on the oracle: returns per sample
    (binary histogram over [-2N, 2N], found lattices, their energies)."""
    N = X * Y
    o = oracle()
    iv = engine.interval_results(-2 * N, 2 * N, num_intervals, 1, 1.0)
    total = S * walkers
    lat = np.zeros(total * N, np.int8)
    probs = np.zeros(total, np.float32)
    o.wlo_init_lattice(p(lat), p(probs), X, Y, total, seed, walkers)
    off = np.zeros(total, np.int32)
    o.wlo_init_offsets_lattice(p(off), X, Y, total)
    E = np.zeros(total, np.int32)
    o.wlo_calc_energy(boundary, p(E), p(lat), p(J), p(off), X, Y, total, walkers)
    hist = np.zeros(S * (4 * N + 1), np.uint64)
    seed_off = np.zeros(total, np.uint64)
    flag = np.zeros(S * num_intervals, np.int32)
    found = np.zeros(S * num_intervals * N, np.int8)
    o.wlo_wang_landau_pre_run(p(E), p(flag), p(lat), p(found), p(hist), p(seed_off), p(J), p(off), -2 * N, 2 * N,
                              iters, X, Y, seed, num_intervals, iv.len_interval, total, walkers, boundary)
    out = []
    for s in range(S):
        h = (hist[s * (4 * N + 1):(s + 1) * (4 * N + 1)] != 0).astype(np.uint8)
        fl = found[s * num_intervals * N:(s + 1) * num_intervals * N].reshape(num_intervals, N)
        keep = fl[:, 0] != 0
        fl = fl[keep]
        fe = engine.lattice_energies(fl, J[s * 2 * N:(s + 1) * 2 * N], X, Y, boundary) if fl.size else np.zeros(0)
        out.append((h, fl.copy(), np.asarray(fe, np.int32)))
    return out


def oracle_samples(cfg: engine.RunConfig, S: int, prob: float, cls: str, prerun_walkers: int = 32,
                   prerun_iters: int = 4000, prerun_intervals: int = 16) -> list[engine.SampleInput]:
    """Disorder + oracle prerun -> one SampleInput per sample (rng key seed + s)."""
    X, Y = cfg.X, cfg.Y
    N = X * Y
    J = oracle_disorder(cfg.seed, S, X, Y, prob, cls)
    pre = oracle_prerun(J, X, Y, cfg.boundary, cfg.seed, S, prerun_walkers, prerun_iters, prerun_intervals)
    samples = []
    for s in range(S):
        h, fl, fe = pre[s]
        emin, emax, mask = engine.energy_spectrum(h, X, Y)
        samples.append(engine.SampleInput(J=J[s * 2 * N:(s + 1) * 2 * N].copy(), e_min=emin, e_max=emax, mask=mask,
                                          found=fl, found_energy=fe, rng_key=cfg.seed + s))
    return samples


def exact_samples(cfg: engine.RunConfig, S: int, prob: float, cls: str) -> list[engine.SampleInput]:
    """Like oracle_samples but with the exact spectrum (transfer matrix) and
    seed lattices from a brute-force search over random lattices: for tests
    that must not depend on the prerun."""
    from oracle_lib import exact_lng  # noqa: PLC0415
    X, Y = cfg.X, cfg.Y
    N = X * Y
    J = oracle_disorder(cfg.seed, S, X, Y, prob, cls)
    pre = oracle_prerun(J, X, Y, cfg.boundary, cfg.seed, S, 32, 4000, 16)
    samples = []
    for s in range(S):
        Js = J[s * 2 * N:(s + 1) * 2 * N].copy()
        lng = exact_lng(Js, X, Y, cfg.boundary)
        h = np.isfinite(lng).astype(np.uint8)
        emin, emax, mask = engine.energy_spectrum(h, X, Y)
        _, fl, fe = pre[s]
        samples.append(engine.SampleInput(J=Js, e_min=emin, e_max=emax, mask=mask, found=fl, found_energy=fe,
                                          rng_key=cfg.seed + s))
    return samples
