"""The C++ `gpu` backend of a FluidSolver (host/gpu_backend.hpp, driven by
host/backend_check the way Project::run_steps calls FluidSolver::step: one
step per call, the state downloaded after each) computes what the Python
binding computes through the same C-ABI: states and step fluxes to rounding
(the two processes may run different launch shapes, whose results agree to
rounding), and a zero density ends the run with the reference's message and
exit status 1 (src/fcalculator_workspace.cc:126-128, src/main.cc:39-42)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import shhd
from helpers import random_active_state, rel_l2
from test_abi import backend_check

pytestmark = pytest.mark.gpu


def model_h(na, precision=64):
    return shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=float(na), ly=float(na), dt=0.005,
                             num_order_parameters=2, pressure_coefficient=1.0, eta=0.5, zeta=0.3,
                             mobility=[1.0, 1.2], free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-1.0, -1.0],
                             mu_k2=[1.0, 1.0], phi4_u=[1.0, 1.0], momentum_advection=True,
                             order_parameter_advection=True, noise_momentum=True, noise_order_parameter=True,
                             kBT=1e-3, seed=17, precision=precision)


def run_backend(cfg, st, nsteps, flux, tmp):
    c = cfg.to_c()
    paths = [os.path.join(tmp, n) for n in ("config.bin", "state.bin", "out.bin")]
    open(paths[0], "wb").write(bytes(c))
    open(paths[1], "wb").write(np.ascontiguousarray(st, dtype=np.complex128).tobytes())
    r = subprocess.run([backend_check(), "run", paths[0], paths[1], str(nsteps), str(int(flux)), paths[2]],
                       capture_output=True, text=True, timeout=120)
    return r, paths[2]


@pytest.mark.parametrize("na,precision,tol", [(64, 64, 1e-12), (512, 64, 1e-12), (512, 32, 1e-5)])
def test_cpp_backend_matches_python_binding(na, precision, tol):
    print(f"[case] C++ backend {na}^2 fp{precision}", flush=True)
    cfg = model_h(na, precision)
    st = random_active_state(cfg, np.random.default_rng(na), amp={0: 0.01, 1: 0.2, 2: 0.2, 3: 0.01, 4: 0.01})
    nsteps = 5
    with tempfile.TemporaryDirectory() as tmp:
        r, out = run_backend(cfg, st, nsteps, True, tmp)
        assert r.returncode == 0, r.stderr
        raw = np.fromfile(out, dtype=np.complex128)
    nst = int(np.prod(cfg.state_shape()))
    got = raw[:nst].reshape(cfg.state_shape())
    got_flux = raw[nst:].reshape(cfg.flux_shape())
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_flux_request(density=True, order_parameter=True, momentum=True)
        g.set_state(st)
        for _ in range(nsteps):
            g.step(1)
        ref = g.get_state()
        ref_flux = g.get_flux()
    for f in range(cfg.num_fields):
        assert rel_l2(got[f], ref[f]) <= tol, f
    for s in range(ref_flux.shape[0]):
        assert rel_l2(got_flux[s], ref_flux[s]) <= tol, ("flux", s)


def test_cpp_backend_reports_zero_density_like_the_reference():
    print("[case] C++ backend zero density", flush=True)
    cfg = shhd.SolverConfig(nx_active=16, ny_active=16, dealias="3/2", lx=16.0, ly=16.0, dt=0.01,
                            pressure_coefficient=1.0, eta=0.1, zeta=0.1, momentum_advection=True)
    st = random_active_state(cfg, np.random.default_rng(5), amp={1: 0.01, 2: 0.01})
    st[0] = 0.0
    with tempfile.TemporaryDirectory() as tmp:
        r, _ = run_backend(cfg, st, 3, False, tmp)
    assert r.returncode == 1
    assert "compressible velocity encountered zero density." in r.stderr
