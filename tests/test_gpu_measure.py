"""On-device FP64 accumulators (shhd_*_set_probes / read_probes,
shhd_*_set_structure_factor / get_structure_factor).

* Probe rows equal, bit for bit, the same modes read from get_state /
  get_flux after every step of a step-by-step run (FP64 and FP32 forms,
  graph replay included), so the measures fed by them see exactly what a
  per-step download would give.
* The structure-factor sums equal the host sums of |f_hat|^2 over the
  sampled states to 1e-14 relative.
* A step that would overflow the ring fails before any work.
"""
import dataclasses

import numpy as np
import pytest

import shhd
import measures
from helpers import flux_configs, flux_state, rel_l2

pytestmark = pytest.mark.gpu

CONFIGS = flux_configs()


def banner(name):
    print(f"[case] {name}", flush=True)


def probe_list(cfg):
    n = cfg.num_order_parameters
    probes = [(shhd.PROBE_STATE, f, gx, gy) for f in (0, 1, cfg.num_fields - 1)
              for gx, gy in [(0, 0)] + measures.lowest_modes(cfg, 5)]
    probes += measures.green_kubo_probes(cfg, "diagonal")
    probes += measures.flux_time_series_probes(cfg, ["pi_xx", "pi_xy", "pi_yy"] +
                                               [f"op_flux_x[{a}]" for a in range(n)])
    return probes


def extract(state, flux, probes):
    return np.array([(state if src == shhd.PROBE_STATE else flux)[idx, gy, gx] for src, idx, gx, gy in probes])


@pytest.mark.parametrize("graphs", ["0", "1"])
@pytest.mark.parametrize("precision", [64, 32])
def test_probe_rows_equal_per_step_downloads(precision, graphs, monkeypatch):
    monkeypatch.setenv("SHHD_USE_GRAPH", graphs)
    cfg = dataclasses.replace(CONFIGS["model_h_srk3"], precision=precision)
    banner(f"probes fp{precision} graphs={graphs}")
    probes = probe_list(cfg)
    st = flux_state(cfg)
    expect = []
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        g.set_flux_request(True, True, True)
        for _ in range(9):
            g.step(1)
            expect.append(extract(g.get_state(), g.get_flux(), probes))
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        g.set_flux_request(True, True, True)
        g.set_probes(probes, capacity=16)
        g.step(6)
        a = g.read_probes()
        g.step(3)
        b = g.read_probes()
    rows = np.concatenate([a, b])
    assert rows.shape == (9, len(probes))
    assert np.array_equal(rows, np.array(expect))


def test_probe_ring_overflow_is_refused():
    cfg = CONFIGS["quiescent_euler"]
    banner("probe overflow")
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(flux_state(cfg))
        g.set_probes([(shhd.PROBE_STATE, 1, 1, 0)], capacity=4)
        with pytest.raises(shhd.ShhdError) as e:
            g.step(5)
        assert e.value.status == shhd.ERR_FULL
        g.step(4)
        assert g.read_probes().shape == (4, 1)
        g.step(3)
        assert g.read_probes().shape == (3, 1)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", ["quiescent_srk3", "incompressible_euler", "model_h_srk3"])
def test_structure_factor_sums(name, precision):
    cfg = dataclasses.replace(CONFIGS[name], precision=precision)
    banner(f"structure factor {name} fp{precision}")
    fields = [0, 1, cfg.num_fields - 1]
    st = flux_state(cfg)
    host = np.zeros((len(fields), cfg.ny, cfg.nx // 2 + 1))
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        for _ in range(4):
            g.step(2)
            s = g.get_state()
            host += np.abs(s[fields]) ** 2
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        g.set_structure_factor(fields, nevery=2)
        g.step(8)
        sk, count = g.get_structure_factor()
    assert count == 4
    assert rel_l2(sk, host) <= 1e-14
    assert np.all(sk[:, ~shhd.active_mask(cfg)] == 0)


@pytest.mark.parametrize("graphs", [False, True])
def test_batch_measures_alone_and_mixed(graphs):
    banner(f"batch measures graphs={graphs}")
    names = ["model_h_srk3", "quiescent_srk3", "incompressible_euler"]
    cfgs = [CONFIGS[n] for n in names]
    states = [flux_state(c, seed=20 + i) for i, c in enumerate(cfgs)]

    def arm(b, slot, c):
        b.set_flux_request(slot, True, True, True)
        b.set_probes(slot, probe_list(c), 32)
        b.set_structure_factor(slot, [1], nevery=3)

    alone = []
    for c, s in zip(cfgs, states):
        with shhd.Batch(use_graphs=graphs) as b:
            slot = b.admit(c, s)
            arm(b, slot, c)
            b.step(10)
            alone.append((b.read_probes(slot), b.get_structure_factor(slot)))
    with shhd.Batch(use_graphs=graphs) as b:
        slots = [b.admit(c, s) for c, s in zip(cfgs, states)]
        for slot, c in zip(slots, cfgs):
            arm(b, slot, c)
        b.step(10)
        for i, slot in enumerate(slots):
            rows = b.read_probes(slot)
            sk, cnt = b.get_structure_factor(slot)
            assert rows.shape[0] == 10 and cnt == 3
            assert np.array_equal(rows, alone[i][0]), names[i]
            assert np.array_equal(sk, alone[i][1][0]) and cnt == alone[i][1][1]
