"""The reference-side lowering (include/shhd_reference_lowering.hpp), compiled
against the reference's own model plugins (oracle/Makefile.ref builds them
from /root/reference/src into oracle/_ref/), gives the configuration the
device and the oracle run:

- phi4 / quadratic / null free energy, linear-EOS / null thermodynamics and
  constant transport lower to the coefficients the Python binding writes, and
  the oracle's right-hand side is bit-identical under both;
- u of phi4 is read through Phi4FreeEnergy::physical_chemical_potential
  (src/model_free_energy_phi4.cc:47-55), and the oracle's cubic term equals
  the transform of the reference's own pointwise function on the padded grid;
- noise targets follow FluidSolver (src/fluid_solver.h:79-94): quiescent runs
  drop the momentum noise.

Runs where the reference sources are present (the build host); skipped
elsewhere.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import shhd
import oracle_py as op

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
LIB = os.path.join(ROOT, "oracle", "_ref", "libref_models.so")
MAXN = shhd.MAX_OP


class RefModelSpec(ctypes.Structure):
    _fields_ = [("nx_active", ctypes.c_int), ("ny_active", ctypes.c_int), ("dealias", ctypes.c_int),
                ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("num_order_parameters", ctypes.c_int),
                ("time_evolution", ctypes.c_char * 32), ("dt", ctypes.c_double), ("thermo", ctypes.c_int),
                ("cT", ctypes.c_double), ("free_energy", ctypes.c_int), ("a", ctypes.c_double * MAXN),
                ("b", ctypes.c_double * MAXN), ("u", ctypes.c_double * MAXN), ("eta", ctypes.c_double),
                ("zeta", ctypes.c_double), ("mobility", ctypes.c_double * MAXN),
                ("momentum_advection", ctypes.c_int), ("order_parameter_advection", ctypes.c_int),
                ("noise_momentum", ctypes.c_int), ("noise_order_parameter", ctypes.c_int), ("seed", ctypes.c_int),
                ("kBT", ctypes.c_double), ("chi", ctypes.c_double)]


@pytest.fixture(scope="module")
def ref():
    if not os.path.isdir(os.path.join(REF, "src")):
        pytest.skip("reference sources not present on this host")
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "Makefile.ref")], check=True)
    lib = ctypes.CDLL(LIB)
    lib.ref_lower.argtypes = [ctypes.POINTER(RefModelSpec), ctypes.POINTER(shhd.ShhdConfig)]
    lib.ref_lower.restype = ctypes.c_int
    lib.ref_last_error.restype = ctypes.c_char_p
    lib.ref_physical_chemical_potential.argtypes = [ctypes.POINTER(RefModelSpec), ctypes.c_int,
                                                    ctypes.POINTER(ctypes.c_double), ctypes.c_size_t,
                                                    ctypes.POINTER(ctypes.c_double)]
    lib.ref_physical_chemical_potential.restype = ctypes.c_int
    return lib


def spec(n=2, evolution="srk3/compressible", fe=2, thermo=1, a=(-1.0, -0.5), b=(1.0, 0.8), u=(1.0, 0.3), **kw):
    s = RefModelSpec()
    s.nx_active, s.ny_active, s.dealias = kw.get("na", 16), kw.get("na", 16), 1
    s.lx, s.ly = 16.0, 13.0
    s.num_order_parameters = n
    s.time_evolution = evolution.encode()
    s.dt = 0.005
    s.thermo, s.cT = thermo, 1.3
    s.free_energy = fe
    for i in range(n):
        s.a[i], s.b[i], s.u[i] = a[i], b[i], u[i]
        s.mobility[i] = 0.9 + 0.1 * i
    s.eta, s.zeta = 0.5, 0.3
    s.momentum_advection = int(kw.get("madv", evolution != "srk3/quiescent"))
    s.order_parameter_advection = int(kw.get("padv", evolution != "srk3/quiescent" and n > 0))
    s.noise_momentum, s.noise_order_parameter = 1, 1
    s.seed, s.kBT, s.chi = 77, 1e-3, 1.0
    return s


def lower(lib, s):
    c = shhd.ShhdConfig()
    if lib.ref_lower(ctypes.byref(s), ctypes.byref(c)) != 0:
        raise RuntimeError(lib.ref_last_error().decode())
    return c


def python_lowered(s):
    """The same input script through the Python binding's lowering
    (phi4 keys a/b/u -> k0/k2/u, quadratic a -> k0, linear_eos cT -> cT^2)."""
    n = s.num_order_parameters
    evo = s.time_evolution.decode()
    mode = {"compressible": 0, "incompressible": 1, "quiescent": 2}[evo.split("/")[1]]
    kind = {0: shhd.FREE_ENERGY_NULL, 1: shhd.FREE_ENERGY_QUADRATIC, 2: shhd.FREE_ENERGY_PHI4}[s.free_energy]
    return shhd.SolverConfig(
        nx_active=s.nx_active, ny_active=s.ny_active, dealias="3/2", lx=s.lx, ly=s.ly, num_order_parameters=n,
        mode=mode, integrator=shhd.INTEGRATOR_SRK3, dt=s.dt, pressure_coefficient=s.cT ** 2 if s.thermo else 0.0,
        eta=s.eta, zeta=s.zeta, mobility=list(s.mobility[:n]), free_energy_kind=kind,
        mu_k0=list(s.a[:n]) if s.free_energy else [0.0] * n,
        mu_k2=list(s.b[:n]) if s.free_energy == 2 else [0.0] * n,
        phi4_u=list(s.u[:n]) if s.free_energy == 2 else [0.0] * n,
        momentum_advection=bool(s.momentum_advection), order_parameter_advection=bool(s.order_parameter_advection),
        noise_momentum=bool(s.noise_momentum) and mode != 2, noise_order_parameter=bool(s.noise_order_parameter) and n > 0,
        kBT=s.kBT, order_parameter_noise_chi=s.chi, seed=s.seed)


def oracle_of(cfg, c):
    """An oracle created from the lowered struct c itself (cfg: the same
    configuration's Python view, for the state shapes)."""
    o = op.Oracle(cfg)
    o.lib.oracle_destroy(o.h)
    h = ctypes.c_void_p()
    assert o.lib.oracle_create(ctypes.byref(c), ctypes.byref(h)) == 0
    o.h, o._c = h, c
    return o


def config_fields(c):
    n = c.num_order_parameters
    out = {}
    for name, _ in shhd.ShhdConfig._fields_:
        v = getattr(c, name)
        if name in ("mobility", "mu_k0", "mu_k2", "mu_k4", "phi4_u"):
            v = list(v[:n])
        elif name in ("sine_forces", "gradient_forces"):
            continue
        out[name] = v
    return out


@pytest.mark.parametrize("evolution,fe,thermo,n", [
    ("srk3/compressible", 2, 1, 2), ("srk3/incompressible", 2, 0, 1), ("srk3/quiescent", 2, 0, 1),
    ("euler/compressible", 1, 1, 1), ("srk3/compressible", 0, 1, 0), ("srk3/quiescent", 1, 0, 2),
])
def test_lowering_matches_binding_and_oracle_rhs(ref, evolution, fe, thermo, n):
    s = spec(n=n, evolution=evolution, fe=fe, thermo=thermo)
    c = lower(ref, s)
    py = python_lowered(s)
    if evolution.startswith("euler"):
        py.integrator = shhd.INTEGRATOR_EULER
    want = config_fields(py.to_c())
    got = config_fields(c)
    assert got == want
    if fe == 2:
        assert list(c.phi4_u[:n]) == list(s.u[:n])   # u reached through physical_chemical_potential
    if evolution.endswith("quiescent"):
        assert c.noise_momentum == 0
    # the oracle's right-hand side is the same under both lowerings
    rng = np.random.default_rng(1)
    st = np.zeros(py.state_shape(), complex)
    mask = shhd.active_mask(py)
    for f in range(py.num_fields):
        st[f] = np.where(mask, rng.normal(size=mask.shape) + 1j * rng.normal(size=mask.shape), 0) * 20
    st[0, 0, 0] = py.nx * py.ny
    st = shhd.hermitian_project(st, py)
    assert np.array_equal(op.Oracle(py).rhs_det(st), oracle_of(py, c).rhs_det(st))


def test_oracle_cubic_term_is_the_reference_pointwise_function(ref):
    """Quiescent phi4 with only u: RHS_psi = -M k^2 FFT(mu(psi)) on the padded
    grid, with mu evaluated by Phi4FreeEnergy::physical_chemical_potential."""
    s = spec(n=2, evolution="srk3/quiescent", fe=2, thermo=0, a=(0.0, 0.0), b=(0.0, 0.0), u=(0.7, 1.9), na=32)
    c = lower(ref, s)
    cfg = python_lowered(s)
    rng = np.random.default_rng(2)
    mask = shhd.active_mask(cfg)
    st = np.zeros(cfg.state_shape(), complex)
    nc = cfg.nx * cfg.ny
    for a in range(2):
        st[1 + a] = np.where(mask, rng.normal(size=mask.shape) + 1j * rng.normal(size=mask.shape), 0) * (0.05 * nc)
    st = shhd.hermitian_project(st, cfg)
    st[0, 0, 0] = nc
    got = oracle_of(cfg, c).rhs_det(st)
    # psi on the padded grid (backward transform scaled by 1/(Mx My), src/fourier_transform.h:50)
    psi = np.stack([np.fft.irfft2(st[1 + a], s=(cfg.ny, cfg.nx)) for a in range(2)], axis=-1)
    pts = np.ascontiguousarray(psi.reshape(-1, 2))
    kx, ky, _ = shhd.wavenumbers(cfg)
    k2 = kx ** 2 + ky ** 2
    for a in range(2):
        mu = np.zeros(pts.shape[0])
        assert ref.ref_physical_chemical_potential(ctypes.byref(s), a, pts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                   pts.shape[0], mu.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
        mu_hat = np.fft.rfft2(mu.reshape(cfg.ny, cfg.nx))   # forward transform, unscaled
        want = np.where(mask, -s.mobility[a] * k2 * mu_hat, 0)
        err = np.abs(got[1 + a] - want).max() / np.abs(want).max()
        assert err <= 1e-12, err


def test_refuses_invalid_phi4_and_reports_reference_errors(ref):
    s = spec(n=1, evolution="srk3/compressible", fe=2, u=(-1.0,))
    with pytest.raises(RuntimeError, match="nonnegative u"):
        lower(ref, s)
    s = spec(n=1, evolution="rk4/compressible")
    with pytest.raises(RuntimeError, match="unknown time_evolution"):
        lower(ref, s)
