"""Launch-shape autotuner (shhd_autotune): it returns shapes the planner can
build, records them for later instances of the shape, and a solver that runs
with the tuned shapes computes what the default shapes compute (the tile and
lane counts only change how the modes are split over workgroups)."""
import numpy as np
import pytest

import shhd
from helpers import flux_configs, flux_state, rel_l2

pytestmark = pytest.mark.gpu


def test_autotune_keeps_results():
    cfg = flux_configs()["model_h_srk3"]
    print("[case] autotune model_h_srk3", flush=True)
    st = flux_state(cfg)
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        g.step(6)
        before = g.get_state()
    t = shhd.autotune(cfg, steps=2)
    assert t["inverse_column_tile"] in (1, 2, 4, 8) and t["forward_column_tile"] in (1, 2, 4, 8)
    assert t["row_lanes"] in (64, 256, 512, 1024)
    assert 0 <= t["plane_block"] <= 3 and 0 <= t["inverse_plane_block"] <= 4
    assert all(m > 0 for m in t["ms"])
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        g.step(6)
        after = g.get_state()
    for f in range(cfg.num_fields):
        assert rel_l2(after[f], before[f]) <= 1e-12
