"""Physical fields on the device (shhd_*_get_physical) against the oracle's
backward transform (pinned against the DFT definition), and spectral
snapshots as a state-exchange format: a device state written, read back and
set again is the same state, bit for bit."""
import dataclasses

import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import flux_configs, flux_state, rel_l2

pytestmark = pytest.mark.gpu

CONFIGS = flux_configs()


def banner(name):
    print(f"[case] {name}", flush=True)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", ["model_h_srk3", "incompressible_euler", "quiescent_srk3"])
def test_physical_fields_match_oracle_transform(name, precision):
    cfg = dataclasses.replace(CONFIGS[name], precision=precision)
    banner(f"physical {name} fp{precision}")
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(flux_state(cfg))
        g.step(3)
        st = g.get_state()
        phys = g.get_physical()
    tol = 1e-13 if precision == 64 else 1e-6
    for f in range(cfg.num_fields):
        ref = op.fft2_c2r(st[f], cfg.nx)
        if np.linalg.norm(ref) == 0:
            assert np.abs(phys[f]).max() == 0
            continue
        assert rel_l2(phys[f], ref) <= tol, f


def test_spectral_snapshot_exchange(tmp_path):
    cfg = CONFIGS["model_h_srk3"]
    banner("snapshot exchange")
    p = str(tmp_path / shhd.snapshot_filename("run", 5, 10))
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(flux_state(cfg))
        g.step(5)
        st = g.get_state()
        shhd.write_snapshot(p, shhd.SNAPSHOT_SPECTRAL, st, cfg, 5, 5 * cfg.dt)
    _, step, _, back = shhd.read_snapshot(p)
    assert step == 5 and np.array_equal(back, st)
    # a new handle restarts its Philox counters, so the continued runs are
    # compared without noise: the read-back state evolves exactly like the original
    det = dataclasses.replace(cfg, noise_momentum=False, noise_order_parameter=False)
    with shhd.GpuFluidSolver(det) as a, shhd.GpuFluidSolver(det) as b:
        a.set_state(st)
        b.set_state(back)
        a.step(4)
        b.step(4)
        assert np.array_equal(a.get_state(), b.get_state())
