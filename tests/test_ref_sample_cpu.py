"""CPU test of the reference-build timing used by bench.py's cpu_baseline leg: the reference's own
poisson_SOR_log (oracle/_ref, compiled from its sources) runs a bounded sample in a child process
and ends it with exit(1) at itmax (src/poisson.c:280-284); the sample reports a positive time per
sweep that grows with the grid."""
import numpy as np
import pytest

from oracle import ref_sor_sample

pytestmark = pytest.mark.skipif(not ref_sor_sample.available(), reason="oracle/_ref not built")


def test_reference_sor_sample_times_sweeps():
    rng = np.random.default_rng(3)
    small = ref_sor_sample.time_sweeps(rng.standard_normal((64, 64)), 40, 1.9, 2)
    large = ref_sor_sample.time_sweeps(rng.standard_normal((256, 256)), 40, 1.9, 2)
    assert 0 < small < large


def test_reference_sor_sample_calibrates():
    f = np.zeros((128, 128))
    f[-2, :] = -10.0
    per, sweeps = ref_sor_sample.sample(f, 1.95, 2, sample_seconds=0.5)
    assert per > 0 and 8 <= sweeps <= 40000
