"""The algorithmic byte model of phase 2 (ukhip.flops.oddeven_phase2_bytes)
that bench.py sets beside the measured HBM traffic: which even steps of which
levels run phase 2 (j > 0 only, none in a level of one step), element sizes
by level, and the level-0 form with D diagonal and B formed on load."""
import sys
import os

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "ultimate-kalman_amd", "python"))
from ukhip import flops  # noqa: E402


def test_phase2_units_and_sizes():
    n, nn = 2, 4
    unit64 = 8 * (4 * nn + 2 * n) + 8 * (5 * nn + 2 * n)  # dense reads, writes, FP64
    assert flops.oddeven_phase2_bytes(n, 1) == 0  # one step: no level runs phase 2
    assert flops.oddeven_phase2_bytes(n, 2) == 0  # level 0 has j = 0 only; level 1 has one step
    assert flops.oddeven_phase2_bytes(n, 4) == unit64  # j = 2 at level 0; level 1 (2 steps) has j = 0 only
    # k = 8: level 0 j = 2, 4, 6; level 1 (4 steps) j = 2; level 2 (2 steps) none
    assert flops.oddeven_phase2_bytes(n, 8) == 4 * unit64
    # level 0 in FP32 with D diagonal and B formed from n FP64 scalings, level 1 dense FP64
    unit0 = 4 * (n + 2 * nn + 2 * n) + 8 * n + 4 * (5 * nn + 2 * n)
    assert flops.oddeven_phase2_bytes(n, 8, fp32_levels=1, level0_diag_synth=True) == 3 * unit0 + unit64
