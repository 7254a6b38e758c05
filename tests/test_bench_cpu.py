"""CPU checks of bench.py's contract that need no GPU: it refuses more than one GPU before touching
one, and the committed rocprofv3 summaries it reads for the roofline object parse to the figures
they state."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def load_bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_committed_profiles_parse():
    b = load_bench()
    # traffic per launch: the scaled FETCH_SIZE + WRITE_SIZE figures each summary states
    assert b.read_traffic("sor_fp64_n1024_traffic.txt") == 27756720
    assert b.read_traffic("sor_fp32_refined_n1024_traffic.txt") == 13958017
    for name, waves in (("sor_fp64_n1024_sq_counters.txt", 3856), ("sor_fp32_refined_n1024_sq_counters.txt", 3856)):
        sq = b.read_sq(name)
        assert sq is not None and sq["SQ_WAVES"] == waves
        assert 0 < sq["SQ_WAIT_ANY"] < sq["SQ_WAVE_CYCLES"]
        assert sq["SQ_INSTS_VALU"] / waves > 100


def test_roofline_fraction_is_a_fraction_of_peak():
    # the roofline object prices the bytes one 5-sweep launch must move (24 B per cell FP64) over
    # the launch time against 8 TB/s: at the 10 us launches measured at 1024^2 it is about 0.31
    b = load_bench()
    r = b.roofline(1024 * 1024, 5, 0.010, 24, "sor_fp64_n1024_traffic.txt", "sor_fp64_n1024_sq_counters.txt", 16,
                   "sor_rb_kernel<double,...,S=5>")
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["frac"] - 24 * 1024 * 1024 / 10e-6 / 8e12) < 1e-12 and 0 < r["frac"] <= 1
    assert r["traffic"] == 27756720
