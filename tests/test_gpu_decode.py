"""GPU stitching + log Z (wl_stitch_tasks) against the host implementation
and the oracle (bit-exact), and the physics of the decoded result: normalised
ln g against the exact density of states, and the logical-class decision
against the decision the exact density of states gives."""

from __future__ import annotations

import math

import numpy as np
import pytest

from helpers import exact_samples
from oracle_lib import exact_lng, oracle, p
from qec_pf_wang_landau_amd import engine

pytestmark = pytest.mark.gpu

CLASSES = engine.ERROR_CLASSES


def run_classes(X, Y, boundary, I, W, S, prob, seed=12345, n=1000, c=20):
    cfg = engine.RunConfig(X=X, Y=Y, boundary=boundary, num_intervals=I, walker_per_interval=W, overlap=0.25,
                           alpha=0.8, beta=1e-6, num_iterations=n, exchange_every=c, seed=seed)
    samples = {cls: exact_samples(cfg, S, prob, cls) for cls in CLASSES}
    states = [engine.build_state(samples[cls], cfg) for cls in CLASSES]
    dev = engine.BatchDevice(engine.merge_states(states))
    dev.run()
    return cfg, samples, dev


def test_stitch_on_device_matches_host_and_oracle(cuda):
    cfg, samples, dev = run_classes(5, 5, 1, 5, 4, 3, 0.1)
    got = dev.pull()
    assert (got.arrays["status"] == 1).all()
    betas = np.full(got.tasks.size, 0.7)
    dec = engine.stitch_tasks_gpu(dev, betas)
    for q, part in enumerate(got.split()):
        for s in range(part.S):
            qq = got.parts[q][1] + s
            a, b = part.sample_intervals(s)
            E, lng = engine.stitch(part.sample_logG(s), a, b, cfg.walker_per_interval, cfg.X, cfg.Y)
            assert np.array_equal(dec.E[qq], E) and np.array_equal(dec.lng[qq], lng)
            cap = int(b[-1] - a[0] + 1)
            oE = np.zeros(cap, np.int32)
            og = np.zeros(cap)
            n = oracle().wlo_stitch(p(np.ascontiguousarray(part.sample_logG(s))), p(np.ascontiguousarray(a)),
                                    p(np.ascontiguousarray(b)), cfg.num_intervals, cfg.walker_per_interval, cfg.X,
                                    cfg.Y, p(oE), p(og), cap)
            assert n == E.size and np.array_equal(oE[:n], E) and np.array_equal(og[:n], lng)
            assert dec.log_z[qq] == engine.log_partition(E, lng, 0.7)


def test_decoded_ln_g_against_exact_dos_and_decisions(cuda):
    """6x6 periodic at p = 0.04, 12 samples x 4 classes.  Normalised ln g
    after ln f = 1e-6 is compared with the transfer-matrix DOS: the error
    is the statistical floor of this Wang-Landau configuration (8 walkers per
    interval, flatness 0.8).  The run is deterministic and equal to the CPU
    oracle's, which gives over these 48 tasks a per-task max-abs error of
    0.02-0.20 (largest at the spectrum edges, where g is a few states) and a
    mean-abs error of 0.019, so the bounds are 0.25 max-abs and 0.03 for the
    mean of the per-task mean-abs errors.  The logical class chosen
    from ln Z at the Nishimori temperature must equal the exact choice
    wherever the exact ln Z gap exceeds 0.5."""
    S, prob = 12, 0.04
    cfg, samples, dev = run_classes(6, 6, 0, 10, 8, S, prob)
    got = dev.pull()
    assert (got.arrays["status"] == 1).all()
    bN = engine.nishimori_beta(prob)
    dec = engine.stitch_tasks_gpu(dev, np.full(got.tasks.size, bN))
    lz = dec.log_z.reshape(len(CLASSES), S).T  # [sample][class]
    exact_lz = np.zeros_like(lz)
    worst = 0.0
    means = []
    for ci, cls in enumerate(CLASSES):
        for s in range(S):
            q = ci * S + s
            J = samples[cls][s].J
            ex = exact_lng(J, 6, 6, 0)
            E, g = dec.E[q], dec.lng[q]
            err = np.abs(g - ex[E + 72])
            worst = max(worst, float(err.max()))
            means.append(float(err.mean()))
            fin = np.isfinite(ex)
            Ee = (np.nonzero(fin)[0] - 72).astype(np.int32)
            exact_lz[s, ci] = oracle().wlo_log_partition(p(Ee), p(np.ascontiguousarray(ex[fin])), Ee.size, bN)
            assert set(E.tolist()) == set(Ee.tolist())
            assert abs(math.log(np.exp(g - g.max()).sum()) + g.max() - 36 * math.log(2)) < 1e-10
    assert worst < 0.25, worst
    assert np.mean(means) < 0.03, np.mean(means)
    sorted_exact = np.sort(exact_lz, axis=1)
    clear = (sorted_exact[:, -1] - sorted_exact[:, -2]) > 0.5
    assert clear.sum() >= S // 2
    assert np.array_equal(engine.decide(lz)[clear], engine.decide(exact_lz)[clear])
    # at p = 0.04 the error's own class ('i') wins for almost every sample
    assert np.mean(engine.decide(exact_lz) == 0) >= 0.75
