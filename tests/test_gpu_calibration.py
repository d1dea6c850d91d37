"""The measurement helpers bench.py uses beside the step's kernels:
mhd_calibration_copy (the 1 GiB copy that scales the traffic counters) and
mhd_calibration_mix (the streaming rate of each kernel's read/write mix)."""
import pytest

from magnetohydrodynamics_simulator_amd import Context

pytestmark = pytest.mark.gpu

MIXES = [(1, 1), (2, 1), (2, 2), (3, 2), (7, 2), (6, 12), (12, 6)]


def test_calibration_kernels():
    print("\n[case] calibration copy and stream mixes, 1M doubles", flush=True)
    n = 1 << 20
    with Context(64, 64, 0.1, 0.1, 0.0) as c:
        c.calibration_copy(8 * n)
        for ni, no in MIXES:
            ms = c.calibration_mix(ni, no, n, reps=3)
            assert 0.0 < ms < 100.0, (ni, no, ms)
        with pytest.raises(RuntimeError, match="mixes"):
            c.calibration_mix(4, 4, n)
        with pytest.raises(RuntimeError, match="bad argument"):
            c.calibration_mix(1, 1, 0)
