"""Parity at the benchmark workloads' full sizes.

- Model-H N=2 at 1024^2 (3/2 -> 1536^2), the throughput workload, against the
  FP64 oracle for three deterministic SRK3 steps (noise off: the oracle draws
  the reference's mt19937_64 stream, the device a Philox stream): FP64 to
  1e-10 and the FP32 form to 1e-5 relative L2 per field, with the autotuned
  and with the untuned launch shapes (fused 1536-point transforms).
- the incompressible-turbulence and quiescent models at 1024^2 against the
  oracle in the same way (the row pass's other two fixed field layouts);
- Model-H N=2 and incompressible turbulence at 4096^2 (3/2 -> 6144^2), where
  the oracle needs about a minute a step: the fused 6144-point transforms
  against the runtime-plan transforms (the length-generic build the oracle
  parity tests pin at smaller sizes), two SRK3 steps with noise on (both runs
  draw the same Philox deviates), 1e-12 (FP64) / 1e-5 (FP32); and the
  size-independent properties of the fused run: k = 0 totals conserved, the
  kx = 0 line Hermitian, inactive modes zero, every value finite.

Every case prints its name before touching the GPU.
"""
import dataclasses

import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def banner(name):
    print(f"[case] {name}", flush=True)


def model_h(na, precision, noise):
    """The Model-H N=2 workload of the throughput benchmark (phi4 a = -1,
    kappa = 1, u = 1; eta = zeta = 0.5, M = 1, cT = 1; dt = 0.005 at dx = 1)."""
    kbt = 1e-3 if noise else 0.0
    return shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=float(na), ly=float(na), dt=0.005,
                             num_order_parameters=2, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                             mobility=[1.0, 1.0], free_energy_kind=shhd.FREE_ENERGY_PHI4,
                             mu_k0=[-1.0, -1.0], mu_k2=[1.0, 1.0], phi4_u=[1.0, 1.0],
                             momentum_advection=True, order_parameter_advection=True,
                             noise_momentum=noise, noise_order_parameter=noise, kBT=kbt, seed=2,
                             precision=precision)


def synthetic_state(cfg, seed=2):
    """rho = 1 + 1e-3 fluctuations, psi ~ 1e-2, momentum ~ 1e-3 (physical rms),
    active modes only, Hermitian kx = 0 line."""
    rng = np.random.default_rng(seed)
    mask = shhd.active_mask(cfg)
    nc = cfg.nx * cfg.ny
    st = np.zeros(cfg.state_shape(), np.complex128)
    for f in range(cfg.num_fields):
        amp = 0.01 if 1 <= f <= cfg.num_order_parameters else 1e-3
        z = rng.standard_normal(mask.shape) + 1j * rng.standard_normal(mask.shape)
        st[f] = np.where(mask, z, 0) * (amp * nc / np.sqrt(2.0 * mask.sum()))
    st[:, 0, 0] = 0
    st[0, 0, 0] = nc
    if cfg.mode == shhd.MODE_INCOMPRESSIBLE:   # transverse momentum
        kx, ky, _ = shhd.wavenumbers(cfg)
        k2 = kx ** 2 + ky ** 2
        n = cfg.num_order_parameters
        with np.errstate(invalid="ignore", divide="ignore"):
            tr = np.where(k2 > 0, (-ky * st[n + 1] + kx * st[n + 2]) / k2, 0)
        st[n + 1] = -ky * tr
        st[n + 2] = kx * tr
    return shhd.hermitian_project(st, cfg)


def run_gpu(cfg, shapes, state=None, ics=(), steps=2):
    shhd.set_launch_shapes(cfg, *shapes)
    with shhd.GpuFluidSolver(cfg) as g:
        if state is not None:
            g.set_state(state)
        for ic in ics:
            g.apply_initial_condition(ic)
        first = g.get_state() if state is None else state
        g.step(steps)
        return first, g.get_state(), g.info()


def check_properties(cfg, first, got, tol_totals):
    assert np.all(np.isfinite(got))
    mask = shhd.active_mask(cfg)
    assert np.all(got[:, ~mask] == 0), "inactive modes must stay zero"
    for gy in range(1, cfg.ny_active // 2):
        assert np.array_equal(got[:, cfg.ny - gy, 0], np.conj(got[:, gy, 0])), f"kx = 0 pair {gy}"
    assert np.all(got[:, 0, 0].imag == 0)
    scale = np.abs(first[:, 0, 0]).max()
    drift = np.abs(got[:, 0, 0] - first[:, 0, 0]).max() / scale
    print(f"    k = 0 totals drift {drift:.1e}", flush=True)
    assert drift <= tol_totals


@pytest.fixture(scope="module")
def oracle_1024():
    cfg = model_h(1024, 64, noise=False)
    st = synthetic_state(cfg)
    o = op.Oracle(cfg)
    o.set_state(st)
    o.step(3, 0.0)
    return st, o.get_state()


# the autotuned shapes of this workload and the untuned defaults (fused
# 1536-point column and row transforms in both)
@pytest.mark.parametrize("precision,shapes,tol", [
    (64, (2, 1, 256, 4, 4, 1), 1e-10),
    (64, (0, 0, 0, 4, 4, 2), 1e-10),
    (32, (2, 1, 256, 3, 4, 2), 1e-5),
])
def test_model_h_1024_against_oracle(oracle_1024, precision, shapes, tol):
    banner(f"model-H N=2 1024^2 vs oracle fp{precision} shapes {shapes}")
    st, ref = oracle_1024
    cfg = model_h(1024, precision, noise=False)
    _, got, info = run_gpu(cfg, shapes, state=st, steps=3)
    assert info.forward_column_fused > 0 and info.row_fused > 0 and info.inverse_column_fused > 0
    errs = [rel_l2(got[f], ref[f]) for f in range(cfg.num_fields)]
    print(f"    per-field rel L2 {['%.2e' % e for e in errs]}", flush=True)
    for f, e in enumerate(errs):
        assert e <= tol, (f, e)
    check_properties(cfg, st, got, 1e-12 if precision == 64 else 1e-6)


def other_layout_1024(name, precision):
    """The incompressible-turbulence and quiescent-ensemble models at 1024^2,
    noise off: the row pass's other two fixed field layouts (jx, jy; psi)."""
    if name == "incompressible":
        return shhd.SolverConfig(nx_active=1024, ny_active=1024, dealias="3/2", dt=1e-4,
                                 mode=shhd.MODE_INCOMPRESSIBLE, eta=1e-4, momentum_advection=True,
                                 noise_momentum=False, kBT=0.0, seed=3, precision=precision)
    return shhd.SolverConfig(nx_active=1024, ny_active=1024, dealias="3/2", lx=1536.0, ly=1536.0, dt=0.02,
                             mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[1.0],
                             free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-0.5], mu_k2=[1.0], phi4_u=[1.0],
                             noise_order_parameter=False, kBT=0.0, seed=5, precision=precision)


@pytest.mark.parametrize("name", ["incompressible", "quiescent"])
def test_other_layouts_1024_against_oracle(name):
    banner(f"{name} 1024^2 vs oracle, fp64 and fp32, untuned shapes")
    cfg64 = other_layout_1024(name, 64)
    st = synthetic_state(cfg64, seed=4)
    o = op.Oracle(cfg64)
    o.set_state(st)
    o.step(3, 0.0)
    ref = o.get_state()
    for precision, tol in ((64, 1e-10), (32, 1e-5)):
        cfg = other_layout_1024(name, precision)
        _, got, info = run_gpu(cfg, (0, 0, 0, 4, 4, 2), state=st, steps=3)
        # the fused row pass of this layout and the fused forward columns (the
        # untuned FP32 inverse-column tile of 4 columns has no fused build)
        assert info.forward_column_fused > 0 and info.row_fused > 0
        print(f"    fp{precision} inverse columns fused: {info.inverse_column_fused}", flush=True)
        errs = [rel_l2(got[f], ref[f]) for f in range(cfg.num_fields) if np.linalg.norm(ref[f]) > 0]
        print(f"    fp{precision} per-field rel L2 {['%.2e' % e for e in errs]}", flush=True)
        assert max(errs) <= tol
        check_properties(cfg, st, got, 1e-12 if precision == 64 else 1e-6)


@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 1e-5)])
def test_model_h_rectangular_against_oracle(precision, tol):
    """Model-H N=2 on a 1024 x 256 grid (rows of 1536 points, columns of 384:
    both fused, on unequal sides) against the oracle, three deterministic
    SRK3 steps, untuned shapes."""
    banner(f"model-H N=2 1024x256 vs oracle fp{precision}")
    def cfg_of(prec):
        return dataclasses.replace(model_h(1024, prec, noise=False), ny_active=256, ly=256.0)
    cfg64 = cfg_of(64)
    st = synthetic_state(cfg64, seed=6)
    o = op.Oracle(cfg64)
    o.set_state(st)
    o.step(3, 0.0)
    ref = o.get_state()
    cfg = cfg_of(precision)
    _, got, info = run_gpu(cfg, (0, 0, 0, 4, 4, 2), state=st, steps=3)
    assert info.forward_column_fused > 0 and info.row_fused > 0
    errs = [rel_l2(got[f], ref[f]) for f in range(cfg.num_fields)]
    print(f"    per-field rel L2 {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= tol
    check_properties(cfg, st, got, 1e-12 if precision == 64 else 1e-6)


def model_h_4096_ics():
    IC = shhd.initial_condition
    return ([IC(shhd.IC_EQUILIBRIUM_DENSITY, seed=2, kBT=1e-3, value=1.0, pressure_coefficient=1.0)]
            + [IC(shhd.IC_EQUILIBRIUM_ORDER_PARAMETER, component=a, seed=2, kBT=1e-4, k0=1.0) for a in (0, 1)]
            + [IC(shhd.IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE, component=d, seed=2, kBT=1e-3) for d in (0, 1)])


def incompressible_4096(precision):
    return shhd.SolverConfig(nx_active=4096, ny_active=4096, dealias="3/2", dt=1e-4,
                             mode=shhd.MODE_INCOMPRESSIBLE, eta=1e-4, momentum_advection=True,
                             noise_momentum=True, kBT=1e-8, seed=3, precision=precision)


def incompressible_4096_ics():
    IC = shhd.initial_condition
    return ([IC(shhd.IC_EQUILIBRIUM_DENSITY, seed=3, kBT=0.0, value=1.0, pressure_coefficient=1.0)]
            + [IC(shhd.IC_SPECTRUM_K4, component=d, seed=3, n0=8.0, urms=1.0) for d in (0, 1)])


# (workload, precision, fused shapes = the autotuned ones, tolerance)
CASES_4096 = [
    ("model_h", 64, (1, 1, 1024, 3, 3, 2), 1e-12),
    ("model_h", 32, (1, 1, 1024, 4, 4, 2), 1e-5),
    ("incompressible", 64, (1, 1, 1024, 3, 3, 1), 1e-12),
    ("incompressible", 32, (1, 1, 1024, 4, 3, 2), 1e-5),
]


@pytest.mark.parametrize("name,precision,shapes,tol", CASES_4096)
def test_4096_fused_against_runtime_plan(name, precision, shapes, tol):
    banner(f"{name} 4096^2 fused vs runtime plan fp{precision} shapes {shapes}")
    if name == "model_h":
        cfg, ics = model_h(4096, precision, noise=True), model_h_4096_ics()
    else:
        cfg, ics = incompressible_4096(precision), incompressible_4096_ics()
    first, got, info = run_gpu(cfg, shapes, ics=ics)
    assert info.forward_column_fused > 0 and info.row_fused > 0 and info.inverse_column_fused > 0
    check_properties(cfg, first, got, 1e-12 if precision == 64 else 1e-6)
    generic = dataclasses.replace(cfg)
    first_g, ref, info_g = run_gpu(generic, shapes[:3] + (1, 1, shapes[5]), ics=ics)
    assert (info_g.inverse_column_fused, info_g.forward_column_fused, info_g.row_fused) == (0, 0, 0)
    assert np.array_equal(first, first_g)
    errs = [rel_l2(got[f], ref[f]) for f in range(cfg.num_fields) if np.linalg.norm(ref[f]) > 0]
    print(f"    per-field rel L2 {['%.2e' % e for e in errs]}", flush=True)
    assert max(errs) <= tol


@pytest.mark.parametrize("name,precision,shapes", [("model_h", 64, (1, 1, 1024, 3, 3, 2)),
                                                   ("model_h", 32, (1, 1, 1024, 4, 4, 2)),
                                                   ("incompressible", 32, (1, 1, 1024, 4, 3, 2))])
def test_4096_inverse_plane_blocks_bitwise(name, precision, shapes):
    """The 6144-point fused builds with kx-blocked column-to-row planes (the
    layout the autotuner picks for the FP32 form at 4096^2) give the states
    of row-major planes bit for bit: only where K1 stores and K2 finds a value
    moves."""
    banner(f"{name} 4096^2 inverse plane blocks fp{precision}")
    if name == "model_h":
        cfg, ics = model_h(4096, precision, noise=True), model_h_4096_ics()
    else:
        cfg, ics = incompressible_4096(precision), incompressible_4096_ics()
    _, ref, _ = run_gpu(cfg, shapes + (0,), ics=ics)
    for tb in (1, 3):
        _, got, info = run_gpu(cfg, shapes + (tb,), ics=ics)
        assert info.row_fused > 0 and info.inverse_column_fused > 0
        assert np.array_equal(got, ref), tb
