"""The 3-D run configuration as a static hierarchy for parity tests.

C5: 3-D Euler radial bump (rho = p = 2 for r < 0.2 about the centre of the
unit cube, 1 outside, at rest), extrapolation at every side, level 1 64^3
in eight 32^3 grids, two finer levels at ratio 2 from the flagged initial
data
(amrclaw_amd.workload.radial_euler3_hierarchy, which the 3-D AMR benchmark
uses too).
"""
from amrclaw_amd.workload import radial_euler3_hierarchy


def config_hierarchy3():
    return radial_euler3_hierarchy()
