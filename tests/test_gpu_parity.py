"""Parity of the HIP path (through the C-ABI) against the FP64 CPU oracle.

Deterministic runs (kBT = 0) must agree field by field to <= 1e-10 relative L2
after 100 steps in FP64; the FP32 form to the per-case tolerance written in
each test.  Every case prints its name before touching the GPU.
"""
import math

import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import rel_l2, random_active_state, sine_momentum_state

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10


def banner(name):
    print(f"[case] {name}", flush=True)


def run_both(cfg, state, nsteps, t0=0.0):
    o = op.Oracle(cfg)
    o.set_state(state)
    o.step(nsteps, t0)
    ref = o.get_state()
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(state)
        g.step(nsteps, t0)
        got = g.get_state()
    return got, ref


def per_field(got, ref):
    return [rel_l2(got[f], ref[f]) for f in range(ref.shape[0])]


def assert_close(got, ref, tol, cfg, skip_zero=True):
    errs = per_field(got, ref)
    total = rel_l2(got, ref)
    print(f"    per-field rel L2 {['%.2e' % e for e in errs]}  total {total:.2e}", flush=True)
    assert np.all(np.isfinite(got))
    for f, e in enumerate(errs):
        if skip_zero and np.linalg.norm(ref[f]) == 0:
            assert np.linalg.norm(got[f]) == 0, f"field {f} should stay zero"
            continue
        assert e <= tol, f"field {f}: rel L2 {e:.3e} > {tol:.1e}"
    # inactive modes are zero, kx = 0 line Hermitian
    mask = shhd.active_mask(cfg)
    assert np.all(got[:, ~mask] == 0)


def check_case(cfg, state, nsteps, tol=FP64_TOL, t0=0.0):
    got, ref = run_both(cfg, state, nsteps, t0)
    assert_close(got, ref, tol, cfg)
    return got, ref


# ---------------------------------------------------------------------------
# compressible
# ---------------------------------------------------------------------------
def test_shear_wave_compressible_srk3():
    """Compressible-NS kT=0 shear wave at 16^2: exact decay, z = -eta k^2 dt."""
    banner("shear wave compressible srk3 16^2 3/2")
    cfg = shhd.SolverConfig(nx_active=16, ny_active=16, dealias="3/2", lx=16.0, ly=16.0, dt=0.05,
                            pressure_coefficient=1.0, eta=0.1, zeta=0.1, momentum_advection=True)
    st = sine_momentum_state(cfg, 1e-3)
    got, ref = check_case(cfg, st, 100)
    k = 2 * math.pi / 16
    z = -0.1 * k * k * 0.05
    R = 1 + z + z * z / 2 + z ** 3 / 6
    assert abs(got[1, 1, 0] / st[1, 1, 0] - R ** 100) < 1e-12


def test_compressible_ns_random_128():
    """Compressible NS, N=0, at 128^2 -> 192^2 (dx = 1), kT=0."""
    banner("compressible NS random 128^2 3/2 srk3 100 steps")
    cfg = shhd.SolverConfig(nx_active=128, ny_active=128, dealias="3/2", lx=128.0, ly=128.0, dt=0.05,
                            pressure_coefficient=1.0, eta=0.1, zeta=0.1, momentum_advection=True)
    rng = np.random.default_rng(11)
    st = random_active_state(cfg, rng, amp={0: 0.01, 1: 0.01, 2: 0.01})
    check_case(cfg, st, 100)


def test_model_h_n2_srk3():
    """Model-H physics (2 phi4 fields) at 32^2 with a stable dt."""
    banner("model-H N=2 32^2 3/2 srk3 100 steps")
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=32.0, ly=32.0, dt=0.005,
                            num_order_parameters=2, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                            mobility=[1.0, 1.0], free_energy_kind=shhd.FREE_ENERGY_PHI4,
                            mu_k0=[-1.0, -1.0], mu_k2=[1.0, 1.0], phi4_u=[1.0, 1.0],
                            momentum_advection=True, order_parameter_advection=True)
    rng = np.random.default_rng(2)
    st = random_active_state(cfg, rng, amp={0: 0.01, 1: 0.3, 2: 0.3, 3: 0.02, 4: 0.02})
    check_case(cfg, st, 100)


@pytest.mark.parametrize("precision,tol", [(64, FP64_TOL), (32, 1e-5)])
def test_model_h_512_default_fused_builds(precision, tol):
    """Model-H N=2 at 512^2 -> 768^2 with the launch shapes of a shape never
    tuned: fused column builds (768-point columns) and the fused row build
    with the compressible field layout (768-point rows, two rows per
    workgroup), against the oracle; 20 SRK3 steps."""
    banner(f"model-H N=2 512^2 default fused builds fp{precision}")
    cfg = shhd.SolverConfig(nx_active=512, ny_active=512, dealias="3/2", lx=512.0, ly=512.0, dt=0.005,
                            num_order_parameters=2, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                            mobility=[1.0, 1.0], free_energy_kind=shhd.FREE_ENERGY_PHI4,
                            mu_k0=[-1.0, -1.0], mu_k2=[1.0, 1.0], phi4_u=[1.0, 1.0],
                            momentum_advection=True, order_parameter_advection=True, precision=precision)
    with shhd.GpuFluidSolver(cfg) as g:
        info = g.info()
    assert info.inverse_column_fused >= 1 and info.forward_column_fused >= 1 and info.row_fused >= 1
    rng = np.random.default_rng(5)
    st = random_active_state(cfg, rng, amp={0: 0.01, 1: 0.3, 2: 0.3, 3: 0.02, 4: 0.02})
    check_case(cfg, st, 20, tol=tol)


def test_model_h_n1_euler_rect():
    """Euler, rectangular grid, Lx != Ly, one order parameter, quadratic + advection."""
    banner("model-H N=1 euler 32x16 3/2 rectangular")
    cfg = shhd.SolverConfig(nx_active=32, ny_active=16, dealias="3/2", lx=20.0, ly=12.0, dt=0.002,
                            num_order_parameters=1, integrator=shhd.INTEGRATOR_EULER,
                            pressure_coefficient=0.5, eta=0.3, zeta=0.1, mobility=[0.7],
                            free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[0.5], mu_k2=[0.5], phi4_u=[2.0],
                            momentum_advection=True, order_parameter_advection=True)
    rng = np.random.default_rng(5)
    st = random_active_state(cfg, rng, amp={0: 0.02, 1: 0.2, 2: 0.02, 3: 0.02})
    check_case(cfg, st, 100)


def test_compressible_dealias_none_and_two():
    banner("compressible dealias none 32^2 and two 16^2")
    for rule, na in (("none", 32), ("2", 16)):
        cfg = shhd.SolverConfig(nx_active=na, ny_active=na, dealias=rule, lx=float(na), ly=float(na), dt=0.02,
                                pressure_coefficient=1.0, eta=0.2, zeta=0.05, momentum_advection=True)
        rng = np.random.default_rng(7)
        st = random_active_state(cfg, rng, amp={0: 0.01, 1: 0.01, 2: 0.01})
        check_case(cfg, st, 50)


def test_compressible_forces():
    """force/sine on momentum and psi, force/gradient (compressible: extra product)."""
    banner("compressible forces N=1 32^2")
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=32.0, ly=32.0, dt=0.01,
                            num_order_parameters=1, pressure_coefficient=1.0, eta=0.2, zeta=0.2,
                            mobility=[1.0], free_energy_kind=shhd.FREE_ENERGY_QUADRATIC, mu_k0=[0.5],
                            momentum_advection=True, order_parameter_advection=True,
                            sine_forces=[dict(momentum=1, component=0, axis=1, nk=2, amplitude=1e-3),
                                         dict(momentum=1, component=1, axis=0, nk=3, amplitude=2e-3),
                                         dict(momentum=0, component=0, axis=1, nk=1, amplitude=1e-2)],
                            gradient_forces=[dict(component=0, direction=0, amplitude=0.3),
                                             dict(component=0, direction=1, amplitude=-0.1)])
    rng = np.random.default_rng(9)
    st = random_active_state(cfg, rng, amp={0: 0.01, 1: 0.1, 2: 0.01, 3: 0.01})
    check_case(cfg, st, 100)


# ---------------------------------------------------------------------------
# incompressible
# ---------------------------------------------------------------------------
def transverse(st, cfg):
    """Project momentum onto k-perp so the incompressible IC check passes
    (src/initial_condition.cc:7-40)."""
    kx, ky, _ = shhd.wavenumbers(cfg)
    k2 = kx ** 2 + ky ** 2
    n = cfg.num_order_parameters
    jx, jy = st[n + 1].copy(), st[n + 2].copy()
    with np.errstate(invalid="ignore", divide="ignore"):
        tr = np.where(k2 > 0, (-ky * jx + kx * jy) / k2, 0)
    st[n + 1] = np.where(k2 > 0, -ky * tr, jx)
    st[n + 2] = np.where(k2 > 0, kx * tr, jy)
    return shhd.hermitian_project(st, cfg)


def test_incompressible_turbulence_srk3():
    """Incompressible turbulence physics at 64^2 (L = 2 pi), advection + viscosity."""
    banner("incompressible turbulence 64^2 3/2 srk3 100 steps")
    cfg = shhd.SolverConfig(nx_active=64, ny_active=64, dealias="3/2", dt=1e-3, mode=shhd.MODE_INCOMPRESSIBLE,
                            eta=1e-2, pressure_coefficient=1.0, momentum_advection=True)
    rng = np.random.default_rng(3)
    st = transverse(random_active_state(cfg, rng, amp={1: 1.0, 2: 1.0}), cfg)
    check_case(cfg, st, 100)


def test_incompressible_scalar_euler_gradient():
    banner("incompressible N=1 euler + gradient force 32^2")
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", dt=1e-3, mode=shhd.MODE_INCOMPRESSIBLE,
                            integrator=shhd.INTEGRATOR_EULER, num_order_parameters=1, eta=5e-2,
                            mobility=[0.1], free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-0.5], mu_k2=[0.02],
                            phi4_u=[1.0], momentum_advection=True, order_parameter_advection=True,
                            gradient_forces=[dict(component=0, direction=1, amplitude=0.5)],
                            sine_forces=[dict(momentum=1, component=0, axis=1, nk=1, amplitude=0.1)])
    rng = np.random.default_rng(4)
    st = transverse(random_active_state(cfg, rng, amp={0: 0.0, 1: 0.2, 2: 0.5, 3: 0.5}, rho0=1.3), cfg)
    check_case(cfg, st, 100)


# ---------------------------------------------------------------------------
# quiescent
# ---------------------------------------------------------------------------
def test_quiescent_phi4_srk3():
    """Quiescent phi4 physics at 64^2 with dx = 1.5 (stable)."""
    banner("quiescent phi4 64^2 3/2 srk3 100 steps")
    cfg = shhd.SolverConfig(nx_active=64, ny_active=64, dealias="3/2", lx=96.0, ly=96.0, dt=0.02,
                            mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[1.0],
                            free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-0.5], mu_k2=[1.0], phi4_u=[1.0])
    rng = np.random.default_rng(100)
    st = random_active_state(cfg, rng, amp={1: 0.3})
    check_case(cfg, st, 100)


def test_quiescent_euler_two_fields():
    banner("quiescent N=2 euler 32^2")
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=40.0, ly=40.0, dt=0.01,
                            mode=shhd.MODE_QUIESCENT, integrator=shhd.INTEGRATOR_EULER, num_order_parameters=2,
                            mobility=[1.0, 0.5], free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-0.5, 0.2],
                            mu_k2=[1.0, 0.5], phi4_u=[1.0, 0.0])
    rng = np.random.default_rng(101)
    st = random_active_state(cfg, rng, amp={1: 0.3, 2: 0.2})
    # rho and j are static in quiescent mode but still carried
    st[0, 0, 0] = 1.0 * cfg.nx * cfg.ny
    check_case(cfg, st, 100)


# ---------------------------------------------------------------------------
# FP32 form
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("steps,tol", [(20, 1e-5)])
def test_fp32_model_h(steps, tol):
    """FP32 form against the FP64 oracle: tolerance 1e-5 relative L2 per field
    after 20 SRK3 steps (the stated bound for the FP32 form)."""
    banner("fp32 model-H N=2 32^2")
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=32.0, ly=32.0, dt=0.005,
                            num_order_parameters=2, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                            mobility=[1.0, 1.0], free_energy_kind=shhd.FREE_ENERGY_PHI4,
                            mu_k0=[-1.0, -1.0], mu_k2=[1.0, 1.0], phi4_u=[1.0, 1.0],
                            momentum_advection=True, order_parameter_advection=True, precision=32)
    rng = np.random.default_rng(2)
    st = random_active_state(cfg, rng, amp={0: 0.01, 1: 0.3, 2: 0.3, 3: 0.02, 4: 0.02})
    cfg64 = shhd.SolverConfig(**{**cfg.__dict__, "precision": 64})
    o = op.Oracle(cfg64)
    o.set_state(st)
    o.step(steps)
    ref = o.get_state()
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        g.step(steps)
        got = g.get_state()
    errs = per_field(got, ref)
    print("    fp32 per-field rel L2", ["%.2e" % e for e in errs], flush=True)
    assert max(errs) <= tol
    # conserved totals are kept in FP64: k = 0 modes exact to FP64 roundoff
    for f in range(cfg.num_fields):
        assert abs(got[f, 0, 0] - st[f, 0, 0]) <= 1e-12 * max(1.0, abs(st[f, 0, 0]))


def test_fp32_incompressible_and_quiescent():
    banner("fp32 incompressible 64^2 and quiescent 64^2")
    cfg = shhd.SolverConfig(nx_active=64, ny_active=64, dealias="3/2", dt=1e-3, mode=shhd.MODE_INCOMPRESSIBLE,
                            eta=1e-2, momentum_advection=True, precision=32)
    rng = np.random.default_rng(3)
    st = transverse(random_active_state(cfg, rng, amp={1: 1.0, 2: 1.0}), cfg)
    got, ref = run_both(cfg, st, 20)
    e = max(per_field(got, ref)[1:])
    print(f"    incompressible fp32 rel L2 {e:.2e}", flush=True)
    assert e <= 1e-5
    cfg = shhd.SolverConfig(nx_active=64, ny_active=64, dealias="3/2", lx=96.0, ly=96.0, dt=0.02,
                            mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[1.0],
                            free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-0.5], mu_k2=[1.0], phi4_u=[1.0],
                            precision=32)
    st = random_active_state(cfg, np.random.default_rng(100), amp={1: 0.3})
    got, ref = run_both(cfg, st, 20)
    e = per_field(got, ref)[1]
    print(f"    quiescent fp32 rel L2 {e:.2e}", flush=True)
    assert e <= 1e-5


# ---------------------------------------------------------------------------
# invariants and error paths
# ---------------------------------------------------------------------------
def test_conserved_totals_fp64():
    banner("conserved totals model-H 32^2 fp64")
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=32.0, ly=32.0, dt=0.005,
                            num_order_parameters=2, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                            mobility=[1.0, 1.0], free_energy_kind=shhd.FREE_ENERGY_PHI4,
                            mu_k0=[-1.0, -1.0], mu_k2=[1.0, 1.0], phi4_u=[1.0, 1.0],
                            momentum_advection=True, order_parameter_advection=True)
    st = random_active_state(cfg, np.random.default_rng(12), amp={0: 0.01, 1: 0.3, 2: 0.3, 3: 0.02, 4: 0.02})
    st[1, 0, 0] = 0.1 * cfg.nx * cfg.ny
    st[3, 0, 0] = 0.05 * cfg.nx * cfg.ny
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        g.step(50)
        got = g.get_state()
    for f in range(cfg.num_fields):
        assert abs(got[f, 0, 0] - st[f, 0, 0]) <= 1e-12 * max(1.0, abs(st[f, 0, 0]))


def test_zero_density_error():
    banner("zero density error path")
    cfg = shhd.SolverConfig(nx_active=16, ny_active=16, dealias="3/2", lx=16.0, ly=16.0, dt=0.01,
                            pressure_coefficient=1.0, eta=0.1, zeta=0.1, momentum_advection=True)
    st = np.zeros(cfg.state_shape(), complex)   # uniform density value 0
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        with pytest.raises(shhd.ShhdError, match="zero density"):
            g.step(1)


def test_inactive_mode_state_rejected():
    banner("inactive mode rejected")
    cfg = shhd.SolverConfig(nx_active=16, ny_active=16, dealias="3/2", lx=16.0, ly=16.0)
    st = np.zeros(cfg.state_shape(), complex)
    st[0, 0, 10] = 1.0   # gx = 10 >= Nax/2
    with shhd.GpuFluidSolver(cfg) as g:
        with pytest.raises(shhd.ShhdError, match="inactive"):
            g.set_state(st)
