"""bench.py's readers of the committed profiles, on the CPU: the roofline's
traffic, issue and streaming-ceiling entries of the JSON line come from
these files, so a summary whose lines no longer parse would silently drop
those entries.  The values are checked against what the profiles' own
derivations fix (traffic within a few percent of the algorithmic 40 B and
64 B per particle; the VALU floor the SQ profile states for the M6 pass)."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)  # main() runs only as a script
    return mod


@pytest.mark.parametrize("profile,n", [("pmc_traffic_landau_m6_128x128_1e7.txt", 10_000_000),
                                       ("pmc_traffic_strong_b_256x256_1e8.txt", 100_000_000)])
def test_traffic_lines(bench, profile, n):
    for kernel, per_particle in (("k_seg_f32", 40), ("k_seg_f64", 64)):
        got = bench.pmc_traffic(kernel, profile)
        assert got is not None, (profile, kernel)
        bytes_, path = got
        assert path == os.path.join("profiles", profile)
        assert 1.0 <= bytes_ / (per_particle * n) < 1.06
    assert bench.pmc_traffic("k_vm_step", "pmc_traffic_vm2d_128x16_204800.txt") is not None


def test_issue_lines(bench):
    m6 = bench.issue_counts("k_seg_f32<M6>")
    cic = bench.issue_counts("k_seg_f32<CIC>")
    assert m6 is not None and cic is not None
    valu, lds, clock = m6
    # 5.44 VALU wave instructions per particle, the 432 LDS bytes of 18
    # 16-byte reads and 18 8-byte atomics, a clock at or below the 2.4 GHz peak
    assert 5.0 < valu < 6.0 and lds == 18 * 16 + 18 * 8
    assert 2.0e9 < clock <= bench.CLOCK_HZ
    # the floor the profile states: 1e7 particles, 1024 SIMDs, 4 clocks each
    floor_us = 1e7 * valu / bench.VALU_WAVE_INSTS_PER_S * 1e6
    assert floor_us == pytest.approx(88.5, abs=0.5)
    assert floor_us * bench.CLOCK_HZ / clock == pytest.approx(92.9, abs=0.5)
    assert cic[1] == 2 * 16 + 2 * 8 and cic[0] < valu
    assert bench.issue_counts("k_seg_f64<M6>") is None


def test_stream_ceiling(bench):
    gbs = bench.stream_ceiling()
    assert gbs is not None and 5000.0 < gbs < 8000.0
