"""CPU checks of the C-ABI library: it loads, exports every function the
headers declare, and its host-side functions (interval partition, spectrum,
seed lattices, stitching, log Z) agree with the oracle and the reference."""

from __future__ import annotations

import ctypes as C
import math
import re
from pathlib import Path

import numpy as np
import pytest

from helpers import exact_samples, oracle_disorder
from oracle_lib import exact_lng, oracle, p, ref, run_loop
from qec_pf_wang_landau_amd import _lib, engine

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(wl_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = _lib.lib()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(_lib.SIGNATURES), "binding table misses a declared function"


def test_errors_are_status_codes_not_exceptions():
    lib = _lib.lib()
    st = np.zeros(4, np.int32)
    en = np.zeros(4, np.int32)
    tot = C.c_int64()
    ln = C.c_int()
    rc = lib.wl_interval_results(-16, 16, 4, 8, C.c_float(1.5), st.ctypes.data, en.ctypes.data, C.byref(tot),
                                 C.byref(ln))
    assert rc == 1 and "Overlap is out of bounds." in _lib.last_error()
    with pytest.raises(_lib.WLError):
        engine.interval_results(-16, 16, 64, 8, 1.0)


def test_null_buffers_are_invalid_arguments():
    """Entry points refuse a null buffer they would read or write with status
    1 before touching the device (so this runs without a GPU); calls with
    nothing to do accept null buffers."""
    lib = _lib.lib()
    null = None
    cases = [
        ("wl_calc_energy", (0, null, null, null, null, 4, 4, 16, 16, null)),
        ("wl_init_interactions", (null, 16, 2, 1, C.c_float(0.1), null)),
        ("wl_init_indices", (null, 64, 8, null)),
        ("wl_reset_d_cond", (null, null, 4, C.c_double(1e-6), 8, null)),
        ("wl_check_interactions_finished", (null, null, null, 4, 2, null)),
        ("wl_accept_decisions", (null, null, 10, null, null)),
        ("wl_exact_dos", (null, 4, 4, 0, 1, null, null)),
    ]
    for name, args in cases:
        assert getattr(lib, name)(*args) == 1, name
        assert "null" in _lib.last_error(), (name, _lib.last_error())
    # nothing to do: null buffers are fine
    assert lib.wl_calc_energy(0, null, null, null, null, 4, 4, 0, 16, null) == 0
    assert lib.wl_init_indices(null, 0, 8, null) == 0
    assert lib.wl_exact_dos(null, 4, 4, 0, 0, null, null) == 0
    st = np.zeros(4, np.int32)
    assert lib.wl_interval_results(-16, 16, 4, 8, C.c_float(0.25), st.ctypes.data, null, null, null) == 1
    assert lib.wl_stitch_log_g(null, null, null, 2, 2, 4, 4, null, null, 10, null) == 1
    assert lib.wl_energy_stride(null, 10) == 0
    assert math.isnan(lib.wl_log_partition(null, null, 3, C.c_double(0.5)))


def test_interval_partition_matches_oracle_and_reference():
    r = ref()
    rng = np.random.default_rng(1)
    for _ in range(300):
        e_min = int(rng.integers(-4000, 0))
        e_max = e_min + int(rng.integers(8, 8000))
        I = int(rng.integers(1, 65))
        o = float(np.float32(rng.choice([0.0, 0.25, 0.5, 1.0, rng.uniform(0, 1)])))
        st = np.zeros(I, np.int32)
        en = np.zeros(I, np.int32)
        tot = np.zeros(1, np.int64)
        ln = np.zeros(1, np.int32)
        rc_o = oracle().wlo_interval_results(e_min, e_max, I, 8, o, p(st), p(en), p(tot), p(ln))
        try:
            iv = engine.interval_results(e_min, e_max, I, 8, o)
            ok = True
        except _lib.WLError:
            ok = False
        assert ok == (rc_o == 0)
        if ok:
            assert np.array_equal(iv.start, st) and np.array_equal(iv.end, en)
            assert iv.len_histogram_over_all_walkers == tot[0] and iv.len_interval == ln[0]
        if r is not None:
            st2 = np.zeros(I, np.int32)
            en2 = np.zeros(I, np.int32)
            rc_r = r.wlref_interval_results(e_min, e_max, I, 8, o, p(st2), p(en2), p(tot), p(ln))
            assert (rc_r == 0) == ok
            if ok:
                assert np.array_equal(iv.start, st2) and np.array_equal(iv.end, en2)


def test_energy_spectrum_golden():
    # test_file_io.cpp: binarised histograms {1,0,0,0,1,...}; E_min/E_max from first/last 1
    X = Y = 2
    h = np.zeros(17, np.uint8)
    h[[1, 5, 9]] = 1
    emin, emax, mask = engine.energy_spectrum(h, X, Y)
    assert (emin, emax) == (-7, 1) and mask.tolist() == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    assert engine.energy_stride(mask) == 4
    with pytest.raises(_lib.WLError):
        engine.energy_spectrum(np.zeros(17, np.uint8), X, Y)


def test_seed_lattice_selection():
    N = 4
    found = np.array([[1, 1, 1, 1], [-1, 1, 1, 1], [-1, -1, 1, 1]], np.int8)
    fe = np.array([-8, 0, 4], np.int32)
    out = engine.select_seed_lattices(found, fe, np.array([-8, -2, 2]), np.array([0, 4, 8]), 2, N).reshape(6, N)
    assert (out[0] == found[0]).all() and (out[1] == found[0]).all()   # lowest energy in [-8, 0]
    assert (out[2] == found[1]).all() and (out[4] == found[2]).all()
    with pytest.raises(_lib.WLError):
        engine.select_seed_lattices(found, fe, np.array([-8, 5]), np.array([-6, 7]), 2, N)


def _stitch_oracle(g, start, end, I, W, X, Y):
    cap = int(end[-1] - start[0] + 1)
    oe = np.zeros(cap, np.int32)
    og = np.zeros(cap)
    n = oracle().wlo_stitch(p(np.ascontiguousarray(g)), p(np.ascontiguousarray(start)),
                            p(np.ascontiguousarray(end)), I, W, X, Y, p(oe), p(og), cap)
    return n, oe[:max(n, 0)], og[:max(n, 0)]


@pytest.mark.parametrize("cls,boundary,X,Y,I,W", [("v", 0, 4, 4, 4, 4), ("i", 1, 4, 4, 3, 4),
                                                  ("c", 0, 5, 5, 5, 3)])
def test_stitch_matches_oracle_after_full_run(cls, boundary, X, Y, I, W):
    cfg = engine.RunConfig(X=X, Y=Y, boundary=boundary, num_intervals=I, walker_per_interval=W, beta=1e-4,
                           num_iterations=1000, exchange_every=20, seed=7)
    st = engine.build_state(exact_samples(cfg, 1, 0.15, cls), cfg)
    assert run_loop(st) > 0
    start, end = st.sample_intervals(0)
    E, lng = engine.stitch(st.sample_logG(0), start, end, W, X, Y)
    n, oE, og = _stitch_oracle(st.sample_logG(0), start, end, I, W, X, Y)
    assert n == E.size and np.array_equal(E, oE) and np.array_equal(lng, og)
    m = lng.max()
    assert abs(m + math.log(np.exp(lng - m).sum()) - X * Y * math.log(2)) < 1e-12


def test_stitch_random_blocks_match_oracle():
    """Synthetic ln g blocks with holes exercise every branch of the key search."""
    rng = np.random.default_rng(9)
    for trial in range(200):
        I = int(rng.integers(2, 7))
        W = int(rng.integers(1, 4))
        iv = engine.interval_results(-40, 40, I, W, float(rng.choice([0.25, 0.5, 0.75])))
        g = np.zeros(iv.len_histogram_over_all_walkers)
        idx = 0
        for k in range(I):
            ln = iv.end[k] - iv.start[k] + 1
            vals = np.cumsum(rng.uniform(0.1, 1.0, ln)) * (1 if rng.random() < 0.5 else -1) + rng.uniform(1, 5)
            holes = rng.random(ln) < 0.3
            vals[holes] = 0.0
            for w in range(W):
                g[idx:idx + ln] = vals
                idx += ln
        n, oE, og = _stitch_oracle(g, iv.start, iv.end, I, W, 3, 3)
        if n < 0:
            with pytest.raises(_lib.WLError):
                engine.stitch(g, iv.start, iv.end, W, 3, 3)
            continue
        E, lng = engine.stitch(g, iv.start, iv.end, W, 3, 3)
        assert np.array_equal(E, oE) and np.array_equal(lng, og), trial


def test_log_partition_and_nishimori_match_oracle():
    J = oracle_disorder(8, 1, 4, 4, 0.2, "i")
    lng = exact_lng(J, 4, 4, 0)
    E = (np.nonzero(np.isfinite(lng))[0] - 32).astype(np.int32)
    g = np.ascontiguousarray(lng[np.isfinite(lng)])
    for beta in (0.0, 0.3, 1.1):
        assert engine.log_partition(E, g, beta) == oracle().wlo_log_partition(p(E), p(g), E.size, beta)
    assert abs(engine.log_partition(E, g, 0.0) - 16 * math.log(2)) < 1e-12
    assert engine.nishimori_beta(0.1) == oracle().wlo_nishimori_beta(0.1)
