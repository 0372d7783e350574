"""GPU checks of the set-up kernels the reference tests one by one:
check_energy_ranges (test/test_check_energy_ranges.cu), init_indices
(test/test_init_indices.cu), init_offsets_lattice
(test/test_init_lattice_offsets.cu) and the Fisher-Yates shuffle of the
replica exchange (src/wl.cu:785-808), each on the reference test's shapes and
against the CPU oracle bit for bit."""

from __future__ import annotations

import numpy as np
import pytest

from oracle_lib import oracle, p
from qec_pf_wang_landau_amd import engine
from qec_pf_wang_landau_amd._lib import call
from qec_pf_wang_landau_amd.engine import dptr, to_device, to_host

pytestmark = pytest.mark.gpu

_KEEP: list = []


@pytest.fixture(autouse=True)
def _keep_alive():
    yield
    _KEEP.clear()


def d(a):
    t = to_device(np.ascontiguousarray(a))
    _KEEP.append(t)
    return t


def test_check_energy_ranges_reference_case(cuda):
    """10x10, 100 samples x 15 intervals x 5 walkers, overlap 0.25: energies
    drawn inside their intervals give no flag; then about 95 % of them are
    moved above E_max and the flags must mark exactly those walkers."""
    X = Y = 10
    S, I, W = 100, 15, 5
    total = S * I * W
    iv = engine.interval_results(-2 * X * Y, 2 * X * Y, I, W, 0.25)
    start = np.tile(iv.start, S).astype(np.int32)
    end = np.tile(iv.end, S).astype(np.int32)
    rng = np.random.default_rng(11)
    k = np.arange(total) // W
    E = rng.integers(start[k], end[k] + 1).astype(np.int32)
    for case in ("inside", "outside"):
        if case == "outside":
            move = rng.random(total) >= 0.05
            E = np.where(move, 2 * X * Y + 10, E).astype(np.int32)
        flags = d(np.zeros(total, np.int8))
        call("wl_check_energy_ranges", dptr(flags), dptr(d(E)), dptr(d(start)), dptr(d(end)), total, W, None)
        got = to_host(flags, np.int8)
        ref = np.zeros(total, np.int8)
        oracle().wlo_check_energy_ranges(p(ref), p(E), p(start), p(end), total, W)
        assert np.array_equal(got, ref)
        outside = (E < start[k]) | (E > end[k])
        assert np.array_equal(got != 0, outside)
        assert got.max() == (0 if case == "inside" else 1)


def test_init_indices_reference_case(cuda):
    """100 samples x 10 intervals x 10 walkers: index = walker within its interval."""
    W, total = 10, 100 * 10 * 10
    idx = d(np.full(total, -1, np.int32))
    call("wl_init_indices", dptr(idx), total, W, None)
    got = to_host(idx, np.int32)
    assert np.array_equal(got, np.tile(np.arange(W, dtype=np.int32), total // W))


def test_init_offsets_lattice_reference_case(cuda):
    """4x4, 10 lattices: offset i * 16; and a launch larger than one block."""
    for X, Y, n in ((4, 4, 10), (6, 5, 1000)):
        off = d(np.full(n, -1, np.int32))
        call("wl_init_offsets_lattice", dptr(off), X, Y, n, None)
        assert np.array_equal(to_host(off, np.int32), (np.arange(n) * X * Y).astype(np.int32))


@pytest.mark.parametrize("S,I,W", [(3, 4, 8), (2, 5, 13)])
def test_fisher_yates_matches_oracle_shuffle(cuda, S, I, W):
    """wl_fisher_yates shuffles every interval's index segment with that
    interval's walker-0 exchange stream.  The oracle's replica exchange does
    the same shuffle before its swaps; with every energy outside the
    neighbouring interval no swap (and no exchange draw) follows, so both
    parities of the oracle reproduce the shuffles of all but each sample's
    last interval, bit for bit, counters included.  The last intervals'
    segments must be permutations."""
    seed = 321
    per_sample = I * W
    total = S * per_sample
    rng = np.random.default_rng(5)
    idx0 = np.tile(np.arange(W, dtype=np.int32), S * I)
    # attempt counters in the low 40 bits, a few earlier exchange draws above
    off0 = (rng.integers(0, 1 << 30, total).astype(np.uint64)
            | (rng.integers(0, 50, total).astype(np.uint64) << np.uint64(40)))
    d_idx, d_off = d(idx0), d(off0)
    call("wl_fisher_yates", dptr(d_idx), dptr(d_off), seed, S * I, W, per_sample, None)
    got_idx, got_off = to_host(d_idx, np.int32), to_host(d_off, np.uint64)

    ref_idx, ref_off = idx0.copy(), off0.copy()
    start = np.zeros(S * I, np.int32)
    end = np.zeros(S * I, np.int32)
    E = np.full(total, 1 << 20, np.int32)  # above every interval: no swap is tried
    off_lat = (np.arange(total) * 16).astype(np.int32)
    off_hist = np.zeros(total, np.int32)
    cond_sample = np.zeros(S, np.int32)
    logG = np.zeros(1, np.float64)
    for even in (1, 0):
        oracle().wlo_replica_exchange(p(off_lat), p(E), p(ref_idx), p(ref_off), p(start), p(end), p(off_hist),
                                      p(cond_sample), p(logG), even, seed, I, per_sample, S, 0)
    for k in range(S * I):
        seg = slice(k * W, (k + 1) * W)
        assert sorted(got_idx[seg].tolist()) == list(range(W))
        if k % I != I - 1:
            assert np.array_equal(got_idx[seg], ref_idx[seg]), k
            assert got_off[k * W] == ref_off[k * W], k
        assert np.array_equal(got_off[k * W + 1:(k + 1) * W], off0[k * W + 1:(k + 1) * W])
    assert not np.array_equal(got_idx, idx0)
