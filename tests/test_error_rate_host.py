"""CPU checks of the logical-error-rate aggregation: the class decision with
ties (wl_decide_classes), the Wilson score interval (wl_wilson_interval)
against textbook values and its symmetry, the crossing of two rate curves,
the decode_results executable's argument errors and the parser of its
output."""

from __future__ import annotations

import math
import stat
import subprocess

import numpy as np
import pytest

from qec_pf_wang_landau_amd import _lib, error_rate


def test_decision_and_failure_weights():
    nan, inf = float("nan"), float("inf")
    lz = np.array([
        [5.0, 1.0, 2.0, 3.0],    # i wins
        [1.0, 5.0, 2.0, 3.0],    # v wins
        [4.0, 1.0, 4.0, 3.0],    # i and h tie: a uniform pick misses i half the time
        [4.0, 4.0, 4.0, 1.0],    # three-way tie with i
        [1.0, 4.0, 2.0, 4.0],    # v and c tie above i
        [1.0, nan, 2.0, 3.0],    # undecided
        [1.0, 2.0, inf, 3.0],    # undecided
        [-7.0, -7.0, -7.0, -7.0],
    ])
    dec, fail = error_rate.decide(lz)
    assert dec.tolist() == [0, 1, 0, 0, 1, -1, -1, 0]
    assert fail[:5].tolist() == [0.0, 1.0, 0.5, pytest.approx(2.0 / 3.0), 1.0]
    assert math.isnan(fail[5]) and math.isnan(fail[6])
    assert fail[7] == 0.75


def test_decision_of_no_samples_and_bad_shape():
    dec, fail = error_rate.decide(np.zeros((0, 4)))
    assert dec.size == 0 and fail.size == 0
    with pytest.raises(_lib.WLError):
        error_rate.decide(np.zeros((3, 0)))


def test_wilson_interval_textbook_values():
    z = 1.96
    # 5 of 10: the Wilson interval of Newcombe (1998), table I, is [0.2366, 0.7634]
    lo, hi = error_rate.wilson(5, 10, z)
    assert lo == pytest.approx(0.2366, abs=5e-5) and hi == pytest.approx(0.7634, abs=5e-5)
    # 0 of n: [0, z^2 / (n + z^2)]; n of n: [n / (n + z^2), 1]
    for n in (1, 10, 250):
        lo, hi = error_rate.wilson(0, n, z)
        assert lo == pytest.approx(0.0, abs=1e-15) and hi == pytest.approx(z * z / (n + z * z), rel=1e-14)
        lo, hi = error_rate.wilson(n, n, z)
        assert hi == pytest.approx(1.0, abs=1e-15) and lo == pytest.approx(n / (n + z * z), rel=1e-14)
    # 1 of 29 (Newcombe, table I): [0.0061, 0.1718]
    lo, hi = error_rate.wilson(1, 29, z)
    assert lo == pytest.approx(0.0061, abs=5e-5) and hi == pytest.approx(0.1718, abs=5e-5)


def test_wilson_interval_symmetry_and_contains_estimate():
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = int(rng.integers(1, 5000))
        k = float(rng.integers(0, n + 1)) + (0.5 if rng.random() < 0.3 and n > 1 else 0.0)
        k = min(k, n)
        z = float(rng.uniform(0.5, 3.5))
        lo, hi = error_rate.wilson(k, n, z)
        lo2, hi2 = error_rate.wilson(n - k, n, z)
        assert lo == pytest.approx(1.0 - hi2, abs=1e-12) and hi == pytest.approx(1.0 - lo2, abs=1e-12)
        assert lo <= k / n <= hi
        # wider for larger z
        lo3, hi3 = error_rate.wilson(k, n, z + 0.5)
        assert lo3 <= lo + 1e-15 and hi3 >= hi - 1e-15


def test_wilson_interval_rejects_bad_input():
    for args in ((1, 0, 1.96), (-1, 10, 1.96), (11, 10, 1.96), (1, 10, 0.0)):
        with pytest.raises(_lib.WLError):
            error_rate.wilson(*args)


def test_missing_output_pointers_are_refused():
    """A null output is an argument error (status 1), never a crash."""
    import ctypes as C  # noqa: PLC0415
    z = np.zeros((2, 4))
    dec = np.zeros(2, np.int32)
    assert _lib.lib().wl_decide_classes(z.ctypes.data_as(C.c_void_p), 2, 4, None,
                                      dec.ctypes.data_as(C.c_void_p)) == 1
    assert _lib.lib().wl_decide_classes(None, 0, 4, None, None) == 0
    lo = C.c_double()
    assert _lib.lib().wl_wilson_interval(C.c_double(1.0), 10, C.c_double(1.96), C.byref(lo), None) == 1


def _pt(L, p, rate):
    return error_rate.Point(L, p, "wl", 10, 10, rate * 10, rate, 0.0, 1.0, 1.96, np.zeros((10, 4)),
                            np.zeros(10, np.int32), np.ones((10, 4), np.int32), 0.0)


def test_crossing_of_two_curves():
    ps = [0.06, 0.08, 0.10, 0.12]
    small = [_pt(4, p, r) for p, r in zip(ps, [0.10, 0.15, 0.20, 0.25])]
    large = [_pt(6, p, r) for p, r in zip(ps, [0.05, 0.10, 0.25, 0.40])]
    # differences 0.05, 0.05, -0.05: crossing halfway between 0.08 and 0.10
    assert error_rate.crossing(small, large) == pytest.approx(0.09)
    assert error_rate.crossing(small, small[:]) == 0.06
    apart = [_pt(6, p, r + 0.5) for p, r in zip(ps, [0.10, 0.15, 0.20, 0.25])]
    assert error_rate.crossing(small, apart) is None
    with pytest.raises(ValueError):
        error_rate.crossing(small, large[:3])


def test_decode_results_argument_errors():
    exe = error_rate.DECODE_RESULTS
    if not exe.exists():
        pytest.skip("executables built without HDF5")
    for args in ([], ["--prob", "0.1", "a", "b", "c"], ["--prob", "0.7", "a", "b", "c", "d"], ["--bogus"]):
        p = subprocess.run([str(exe), *args], capture_output=True, text=True)
        assert p.returncode == 1 and "usage" in p.stderr
    p = subprocess.run([str(exe), "--prob", "0.1", "/nonexistent/i.h5", "b", "c", "d"], capture_output=True,
                       text=True)
    assert p.returncode == 1 and "Could not open HDF5 file" in p.stderr
    with pytest.raises(RuntimeError, match="decode_results failed"):
        error_rate.decode_result_files(0.1, ["/nonexistent/i.h5", "b", "c", "d"])


def test_decode_result_files_parses_the_output(tmp_path):
    fake = tmp_path / "fake_decode"
    fake.write_text("#!/bin/sh\n"
                    "echo '0 1.5 0.5 0.25 0.125 i'\n"
                    "echo '3 0.5 1.5 0.25 0.125 v'\n"
                    "echo '4 nan 1 1 1 -'\n"
                    "echo '# samples 2 failures 1 rate 0.5 wilson 0.1 0.9 z 1.96 missing 1'\n")
    fake.chmod(fake.stat().st_mode | stat.S_IEXEC)
    out = error_rate.decode_result_files(0.1, ["a", "b", "c", "d"], binary=fake)
    assert out["disorder"].tolist() == [0, 3, 4]
    assert out["decision"].tolist() == [0, 1, -1]
    assert out["log_z"][1].tolist() == [0.5, 1.5, 0.25, 0.125]
    assert out["samples"] == 2 and out["failures"] == 1 and out["rate"] == 0.5 and out["missing"] == 1
    assert out["wilson"] == [0.1, 0.9]
