"""GPU parity of the batched persistent kernel (wl_fused_run) against the CPU
oracle's mainrun loop: every task's final state (lattices, energies, RNG
counters, ln g, histograms, factors, flags, replica indices and lattice
offsets) must be bit-identical to running that sample alone through the
oracle, for tasks of different shapes sharing one launch."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import exact_samples
from oracle_lib import run_loop
from qec_pf_wang_landau_amd import engine

pytestmark = pytest.mark.gpu


def make_state(X, Y, boundary, I, W, S, prob, cls, n=1000, c=20, seed=12345):
    cfg = engine.RunConfig(X=X, Y=Y, boundary=boundary, num_intervals=I, walker_per_interval=W, overlap=0.25,
                           alpha=0.8, beta=1e-6, num_iterations=n, exchange_every=c, seed=seed)
    return engine.build_state(exact_samples(cfg, S, prob, cls), cfg)


def oracle_per_task(st: engine.MainrunState, max_rounds: int) -> list[engine.MainrunState]:
    """Each sample of st run alone (S = 1) through the oracle loop."""
    out = []
    cfg = st.cfg
    I, W, N = cfg.num_intervals, cfg.walker_per_interval, st.N
    for s in range(st.S):
        one = st.copy()
        w0, nw = s * I * W, I * W
        h0 = int(st.off_hist[w0])
        one.S = 1
        one.J = st.J[s * 2 * N:(s + 1) * 2 * N].copy()
        one.lattice = st.lattice[w0 * N:(w0 + nw) * N].copy()
        one.off_lattice = (st.off_lattice[w0:w0 + nw] - w0 * N).astype(np.int32)
        for k in ("E", "off_iter", "factor", "idx"):
            setattr(one, k, getattr(st, k)[w0:w0 + nw].copy())
        one.logG = st.logG[h0:h0 + st.len_hist[s]].copy()
        one.H = st.H[h0:h0 + st.len_hist[s]].copy()
        one.off_hist = (st.off_hist[w0:w0 + nw] - h0).astype(np.int32)
        one.cond = st.cond[s * I:(s + 1) * I].copy()
        one.cond_sample = st.cond_sample[s:s + 1].copy()
        one.start = st.start[s * I:(s + 1) * I].copy()
        one.end = st.end[s * I:(s + 1) * I].copy()
        sp0 = st.off_spec[s]
        sp1 = st.off_spec[s + 1] if s + 1 < st.S else st.spectrum.size
        one.spectrum = st.spectrum[sp0:sp1].copy()
        one.off_spec = np.zeros(1, np.int32)
        one.len_hist = st.len_hist[s:s + 1].copy()
        one.keys = st.keys[s:s + 1].copy()
        one.counter = np.ones(1, np.int64)
        r = run_loop(one, max_rounds)
        one.rounds = np.array([r], np.int64)
        out.append(one)
    return out


def compare(got: engine.MainrunState, refs: list[engine.MainrunState]):
    cfg = got.cfg
    I, W, N = cfg.num_intervals, cfg.walker_per_interval, got.N
    for s, ref in enumerate(refs):
        w0, nw = s * I * W, I * W
        h0 = int(got.off_hist[w0])
        lat = got.lattice[w0 * N:(w0 + nw) * N]
        assert np.array_equal(lat, ref.lattice), f"task {s} lattice"
        assert np.array_equal(got.off_lattice[w0:w0 + nw] - w0 * N, ref.off_lattice), f"task {s} off_lattice"
        for k in ("E", "off_iter", "factor", "idx"):
            assert np.array_equal(getattr(got, k)[w0:w0 + nw], getattr(ref, k)), f"task {s} {k}"
        assert np.array_equal(got.logG[h0:h0 + got.len_hist[s]], ref.logG), f"task {s} logG"
        assert np.array_equal(got.H[h0:h0 + got.len_hist[s]], ref.H), f"task {s} H"
        assert np.array_equal(got.cond[s * I:(s + 1) * I], ref.cond), f"task {s} cond"
        assert got.cond_sample[s] == ref.cond_sample[0], f"task {s} cond_sample"
        assert got.counter[s] == ref.counter[0], f"task {s} counter"


@pytest.mark.parametrize("max_rounds", [1, 5, 33])
def test_fused_partial_rounds_bit_exact(cuda, max_rounds):
    st = make_state(4, 4, 0, 4, 4, 2, 0.1, "v", c=3)
    refs = oracle_per_task(st, max_rounds)
    b = engine.merge_states([st])
    dev = engine.BatchDevice(b)
    dev.run(max_rounds)
    got = dev.pull().split()[0]
    compare(got, refs)
    assert (got.rounds == [r.rounds[0] for r in refs]).all()


def test_fused_resume_across_launches_bit_exact(cuda):
    st = make_state(4, 4, 1, 3, 4, 2, 0.2, "h", c=2)
    refs = oracle_per_task(st, 40)
    dev = engine.BatchDevice(engine.merge_states([st]))
    for _ in range(4):
        dev.run(10)
    got = dev.pull().split()[0]
    compare(got, refs)


def test_fused_c1_to_convergence_bit_exact(cuda):
    st = make_state(4, 4, 0, 4, 4, 1, 0.1, "v")
    refs = oracle_per_task(st, 1 << 40)
    dev = engine.BatchDevice(engine.merge_states([st]))
    dev.run()
    got = dev.pull().split()[0]
    compare(got, refs)
    assert got.status[0] == 1 and got.cond_sample[0] == -1


def test_fused_uneven_batch_bit_exact(cuda):
    """Tasks of different sizes, boundaries and shapes in one launch."""
    states = [
        make_state(4, 4, 0, 4, 4, 2, 0.1, "i", n=300, c=4),
        make_state(5, 5, 1, 5, 6, 1, 0.15, "v", n=400, c=3),
        make_state(6, 6, 0, 10, 8, 2, 0.04, "c", n=500, c=5),
        make_state(6, 4, 1, 3, 16, 1, 0.2, "h", n=200, c=1),
    ]
    refs = [oracle_per_task(st, 60) for st in states]
    dev = engine.BatchDevice(engine.merge_states(states))
    dev.run(60)
    for got, rf in zip(dev.pull().split(), refs):
        compare(got, rf)


def test_fused_matches_plain_loop(cuda):
    """The batched form and the plain launch sequence agree on a 6x6 batch."""
    st = make_state(6, 6, 0, 10, 8, 3, 0.04, "v", n=1000, c=20)
    ds = engine.DeviceState(st.copy())
    engine.run_plain(ds, 200)
    plain = ds.pull()
    dev = engine.BatchDevice(engine.merge_states([st]))
    dev.run(200)
    fused = dev.pull().split()[0]
    for k in ("lattice", "E", "off_iter", "logG", "H", "factor", "idx", "cond", "cond_sample", "off_lattice"):
        assert np.array_equal(getattr(plain, k), getattr(fused, k)), k
