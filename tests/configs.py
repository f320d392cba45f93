"""The 2-D run configurations as static hierarchies for parity tests
(amrclaw_amd.workload.config_hierarchy, which the configuration benchmarks
use too): C1 advection, periodic, ratios 2, 2; C2 acoustics on [-1, 1]^2,
ratios 2, 4; C3 Euler quadrants over four level-1 grids, ratios 2, 4."""
from amrclaw_amd.workload import CONFIGS, config_hierarchy, config_state  # noqa: F401


def berger_oliger(impl, H, level=1, log=None):
    """One step of `level` and the subcycled steps of the finer levels with
    each level's own ratio (tick.f:230-300 with fixed time steps)."""
    r = impl.advanc(level)
    if log is not None:
        log.append((level, r))
    if level < H.nlev:
        for _ in range(H.levels[level - 1]["ratx"]):
            berger_oliger(impl, H, level + 1, log)
        impl.update(level)
