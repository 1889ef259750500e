"""ctypes binding of libshhd_gpu.so (declared in include/shhd_gpu.h).

`GpuFluidSolver` mirrors the reference's `FluidSolver` (src/fluid_solver.h:23-108):
it is built from the lowered configuration, and `step(state, t)` advances a
reference-layout spectral state by one time step.  State arrays are complex128
numpy arrays of shape (N+3, My, Mx//2+1) with fields ordered
[rho, psi_0..psi_{N-1}, jx, jy] (src/state.cc:8-38).

The library must have been built (`make` in this directory); importing this
module on a machine without it raises immediately — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libshhd_gpu.so")

MAX_OP = 8
MAX_FORCES = 8

MODE_COMPRESSIBLE, MODE_INCOMPRESSIBLE, MODE_QUIESCENT = 0, 1, 2
INTEGRATOR_EULER, INTEGRATOR_SRK3 = 0, 1
FREE_ENERGY_NULL, FREE_ENERGY_QUADRATIC, FREE_ENERGY_PHI4 = 0, 1, 2

STATUS = {0: "SHHD_OK", 1: "SHHD_ERR_INVALID", 2: "SHHD_ERR_UNSUPPORTED", 3: "SHHD_ERR_HIP",
          4: "SHHD_ERR_ZERO_DENSITY", 5: "SHHD_ERR_RUNTIME", 6: "SHHD_ERR_FULL"}
ERR_INVALID = 1
ERR_ZERO_DENSITY = 4
ERR_FULL = 6


class ShhdSineForce(ctypes.Structure):
    _fields_ = [("momentum", ctypes.c_int), ("component", ctypes.c_int), ("axis", ctypes.c_int),
                ("nk", ctypes.c_int), ("amplitude", ctypes.c_double)]


class ShhdProbe(ctypes.Structure):
    _fields_ = [("source", ctypes.c_int), ("index", ctypes.c_int), ("gx", ctypes.c_int), ("gy", ctypes.c_int)]


PROBE_STATE, PROBE_FLUX = 0, 1

(IC_EQUILIBRIUM_DENSITY, IC_EQUILIBRIUM_ORDER_PARAMETER, IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE,
 IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE, IC_RANDOM_VORTICITY, IC_SPECTRUM_K4) = range(6)


class ShhdTuning(ctypes.Structure):
    _fields_ = [("inverse_column_tile", ctypes.c_int), ("forward_column_tile", ctypes.c_int),
                ("row_lanes", ctypes.c_int), ("column_transform", ctypes.c_int), ("row_transform", ctypes.c_int),
                ("plane_block", ctypes.c_int), ("inverse_plane_block", ctypes.c_int),
                ("forward_column_waves", ctypes.c_int), ("row_waves", ctypes.c_int), ("ms", ctypes.c_double * 3)]


class ShhdInitialCondition(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("component", ctypes.c_int), ("seed", ctypes.c_int),
                ("kBT", ctypes.c_double), ("value", ctypes.c_double), ("pressure_coefficient", ctypes.c_double),
                ("k0", ctypes.c_double), ("k2", ctypes.c_double), ("k4", ctypes.c_double),
                ("n0", ctypes.c_double), ("sigma", ctypes.c_double), ("urms", ctypes.c_double)]


def initial_condition(kind, component=0, seed=12345, kBT=0.0, value=0.0, pressure_coefficient=0.0, k0=0.0,
                      k2=0.0, k4=0.0, n0=0.0, sigma=0.0, urms=0.0) -> ShhdInitialCondition:
    """Lowered `set` command (include/shhd_gpu.h shhd_initial_condition); for
    density equilibrium/gaussian pass pressure_coefficient = cT^2."""
    return ShhdInitialCondition(int(kind), int(component), int(seed), float(kBT), float(value),
                                float(pressure_coefficient), float(k0), float(k2), float(k4), float(n0),
                                float(sigma), float(urms))


class ShhdGradientForce(ctypes.Structure):
    _fields_ = [("component", ctypes.c_int), ("direction", ctypes.c_int), ("amplitude", ctypes.c_double)]


class ShhdConfig(ctypes.Structure):
    _fields_ = [
        ("nx_active", ctypes.c_int), ("ny_active", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double),
        ("num_order_parameters", ctypes.c_int),
        ("mode", ctypes.c_int), ("integrator", ctypes.c_int),
        ("dt", ctypes.c_double), ("precision", ctypes.c_int),
        ("pressure_coefficient", ctypes.c_double),
        ("eta", ctypes.c_double), ("zeta", ctypes.c_double),
        ("mobility", ctypes.c_double * MAX_OP),
        ("free_energy_kind", ctypes.c_int),
        ("mu_k0", ctypes.c_double * MAX_OP), ("mu_k2", ctypes.c_double * MAX_OP),
        ("mu_k4", ctypes.c_double * MAX_OP), ("phi4_u", ctypes.c_double * MAX_OP),
        ("momentum_advection", ctypes.c_int), ("order_parameter_advection", ctypes.c_int),
        ("noise_momentum", ctypes.c_int), ("noise_order_parameter", ctypes.c_int),
        ("kBT", ctypes.c_double), ("order_parameter_noise_chi", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("num_sine_forces", ctypes.c_int), ("sine_forces", ShhdSineForce * MAX_FORCES),
        ("num_gradient_forces", ctypes.c_int), ("gradient_forces", ShhdGradientForce * MAX_FORCES),
    ]


class ShhdInfo(ctypes.Structure):
    _fields_ = [("num_fields", ctypes.c_int), ("num_inverse_fields", ctypes.c_int),
                ("num_forward_fields", ctypes.c_int), ("num_updated_fields", ctypes.c_int),
                ("stages_per_step", ctypes.c_int), ("state_doubles", ctypes.c_size_t),
                ("device_bytes", ctypes.c_size_t), ("inverse_column_tile", ctypes.c_int),
                ("forward_column_tile", ctypes.c_int), ("row_lanes", ctypes.c_int),
                ("inverse_column_fused", ctypes.c_int), ("forward_column_fused", ctypes.c_int),
                ("row_fused", ctypes.c_int), ("inverse_column_lanes", ctypes.c_int),
                ("forward_column_lanes", ctypes.c_int)]


@dataclass
class SolverConfig:
    """Plain-Python view of shhd_config with the reference's defaults
    (src/simulationinfo.h:35-124: grid 128^2, L = 2 pi, dt 0.005, srk3/compressible,
    noise seed 12345 and kBT 1)."""
    nx_active: int = 128
    ny_active: int = 128
    dealias: str = "none"           # none | 3/2 | 2 (GridConfig::compute_grid_size)
    lx: float = 2.0 * math.pi
    ly: float = 2.0 * math.pi
    num_order_parameters: int = 0
    mode: int = MODE_COMPRESSIBLE
    integrator: int = INTEGRATOR_SRK3
    dt: float = 0.005
    precision: int = 64
    pressure_coefficient: float = 0.0
    eta: float = 0.0
    zeta: float = 0.0
    mobility: Sequence[float] = ()
    free_energy_kind: int = FREE_ENERGY_NULL
    mu_k0: Sequence[float] = ()
    mu_k2: Sequence[float] = ()
    mu_k4: Sequence[float] = ()
    phi4_u: Sequence[float] = ()
    momentum_advection: bool = False
    order_parameter_advection: bool = False
    noise_momentum: bool = False
    noise_order_parameter: bool = False
    kBT: float = 1.0
    order_parameter_noise_chi: float = 1.0
    seed: int = 12345
    sine_forces: List[dict] = field(default_factory=list)
    gradient_forces: List[dict] = field(default_factory=list)

    @property
    def nx(self) -> int:
        return compute_grid_size(self.nx_active, self.dealias)

    @property
    def ny(self) -> int:
        return compute_grid_size(self.ny_active, self.dealias)

    @property
    def num_fields(self) -> int:
        return self.num_order_parameters + 3

    def state_shape(self):
        return (self.num_fields, self.ny, self.nx // 2 + 1)

    def flux_shape(self):
        """FluxBuffer slots (src/buffer_flux.cc:45-48): 2 + 2N + 3."""
        return (2 + 2 * self.num_order_parameters + 3, self.ny, self.nx // 2 + 1)

    def to_c(self) -> ShhdConfig:
        c = ShhdConfig()
        c.nx_active, c.ny_active = self.nx_active, self.ny_active
        c.nx, c.ny = self.nx, self.ny
        c.lx, c.ly = self.lx, self.ly
        c.num_order_parameters = self.num_order_parameters
        c.mode, c.integrator = self.mode, self.integrator
        c.dt, c.precision = self.dt, self.precision
        c.pressure_coefficient, c.eta, c.zeta = self.pressure_coefficient, self.eta, self.zeta
        n = self.num_order_parameters
        for name in ("mobility", "mu_k0", "mu_k2", "mu_k4", "phi4_u"):
            vals = list(getattr(self, name))
            arr = getattr(c, name)
            for a in range(MAX_OP):
                arr[a] = float(vals[a]) if a < len(vals) else 0.0
            if len(vals) > n and any(v != 0 for v in vals[n:]):
                raise ValueError(f"{name} has more entries than order parameters")
        c.free_energy_kind = self.free_energy_kind
        c.momentum_advection = int(bool(self.momentum_advection))
        c.order_parameter_advection = int(bool(self.order_parameter_advection))
        c.noise_momentum = int(bool(self.noise_momentum))
        c.noise_order_parameter = int(bool(self.noise_order_parameter))
        c.kBT, c.order_parameter_noise_chi = self.kBT, self.order_parameter_noise_chi
        c.seed = int(self.seed)
        c.num_sine_forces = len(self.sine_forces)
        for i, f in enumerate(self.sine_forces):
            c.sine_forces[i] = ShhdSineForce(int(f["momentum"]), int(f["component"]), int(f["axis"]),
                                             int(f["nk"]), float(f["amplitude"]))
        c.num_gradient_forces = len(self.gradient_forces)
        for i, f in enumerate(self.gradient_forces):
            c.gradient_forces[i] = ShhdGradientForce(int(f["component"]), int(f["direction"]),
                                                     float(f["amplitude"]))
        return c


def compute_grid_size(active: int, rule: str) -> int:
    """GridConfig::compute_grid_size (src/simulationinfo.h:44-55)."""
    if active <= 0:
        raise ValueError("grid size must be positive.")
    if rule in ("none", "off"):
        return active
    if rule in ("3/2", "three_halves"):
        if active % 2:
            raise ValueError("3/2 dealiasing requires even active grid size.")
        return 3 * active // 2
    if rule in ("2", "two"):
        return 2 * active
    raise ValueError("unknown dealias rule: " + rule)


# ---------------------------------------------------------------------------
# layout helpers (src/domain.cc:139-149, src/spectral_mask.cc:45-52)
# ---------------------------------------------------------------------------
def wavenumbers(cfg: SolverConfig):
    gx = np.arange(cfg.nx // 2 + 1)
    gy = np.arange(cfg.ny)
    kx = 2.0 * math.pi * gx / cfg.lx
    sy = np.where(gy <= cfg.ny // 2, gy, gy - cfg.ny)
    ky = 2.0 * math.pi * sy / cfg.ly
    return kx[None, :], ky[:, None], sy


def active_mask(cfg: SolverConfig) -> np.ndarray:
    gx = np.arange(cfg.nx // 2 + 1)[None, :]
    gy = np.arange(cfg.ny)[:, None]
    sy = np.where(gy <= cfg.ny // 2, gy, gy - cfg.ny)
    return (gx < cfg.nx_active // 2) & (np.abs(sy) < cfg.ny_active // 2)


def hermitian_project(state: np.ndarray, cfg: SolverConfig) -> np.ndarray:
    """Mask + kx = 0 Hermitian projection (src/time_integrator.h:70-109)."""
    s = np.where(active_mask(cfg)[None], state, 0)
    ny = cfg.ny
    out = s.copy()
    for gy in range(ny):
        cg = (ny - gy) % ny
        if not active_mask(cfg)[gy, 0]:
            continue
        if cg == gy:
            out[:, gy, 0] = out[:, gy, 0].real
        elif gy < cg:
            p = 0.5 * (out[:, gy, 0] + np.conj(out[:, cg, 0]))
            out[:, gy, 0] = p
            out[:, cg, 0] = np.conj(p)
    return out


# ---------------------------------------------------------------------------
# library
# ---------------------------------------------------------------------------
_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libshhd_gpu.so is not built ({path}); run `make` in {HERE}")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    lib.shhd_gpu_create.argtypes = [ctypes.POINTER(ShhdConfig), ctypes.POINTER(P)]
    lib.shhd_gpu_create.restype = ctypes.c_int
    lib.shhd_gpu_set_state.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
    lib.shhd_gpu_set_state.restype = ctypes.c_int
    lib.shhd_gpu_get_state.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
    lib.shhd_gpu_get_state.restype = ctypes.c_int
    lib.shhd_gpu_step.argtypes = [P, ctypes.c_int, ctypes.c_double]
    lib.shhd_gpu_step.restype = ctypes.c_int
    lib.shhd_gpu_set_flux_request.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.shhd_gpu_set_flux_request.restype = ctypes.c_int
    lib.shhd_gpu_get_flux.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
    lib.shhd_gpu_get_flux.restype = ctypes.c_int
    lib.shhd_gpu_destroy.argtypes = [P]
    lib.shhd_gpu_destroy.restype = None
    lib.shhd_gpu_last_error.argtypes = [P]
    lib.shhd_gpu_last_error.restype = ctypes.c_char_p
    lib.shhd_gpu_get_info.argtypes = [P, ctypes.POINTER(ShhdInfo)]
    lib.shhd_gpu_get_info.restype = ctypes.c_int
    lib.shhd_gpu_set_kernel_timing.argtypes = [P, ctypes.c_int]
    lib.shhd_gpu_set_kernel_timing.restype = ctypes.c_int
    lib.shhd_gpu_get_kernel_timing.argtypes = [P, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
    lib.shhd_gpu_get_kernel_timing.restype = ctypes.c_int
    lib.shhd_gpu_get_kernel_bytes.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.shhd_gpu_get_kernel_bytes.restype = ctypes.c_int
    lib.shhd_batch_create.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(P)]
    lib.shhd_batch_create.restype = ctypes.c_int
    lib.shhd_batch_footprint.argtypes = [ctypes.POINTER(ShhdConfig), ctypes.POINTER(ctypes.c_size_t)]
    lib.shhd_batch_footprint.restype = ctypes.c_int
    lib.shhd_batch_admit.argtypes = [P, ctypes.POINTER(ShhdConfig), ctypes.POINTER(ctypes.c_double), ctypes.c_size_t,
                                     ctypes.POINTER(ctypes.c_int)]
    lib.shhd_batch_admit.restype = ctypes.c_int
    lib.shhd_batch_step.argtypes = [P, ctypes.c_int]
    lib.shhd_batch_step.restype = ctypes.c_int
    lib.shhd_batch_hold.argtypes = [P, ctypes.c_int, ctypes.c_int]
    lib.shhd_batch_hold.restype = ctypes.c_int
    lib.shhd_batch_failures.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong),
                                        ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.shhd_batch_failures.restype = ctypes.c_int
    lib.shhd_autotune.argtypes = [ctypes.POINTER(ShhdConfig), ctypes.c_int, ctypes.POINTER(ShhdTuning)]
    lib.shhd_autotune.restype = ctypes.c_int
    lib.shhd_set_tuning.argtypes = [ctypes.POINTER(ShhdConfig), ctypes.POINTER(ShhdTuning)]
    lib.shhd_set_tuning.restype = ctypes.c_int
    lib.shhd_write_snapshot.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    lib.shhd_write_snapshot.restype = ctypes.c_int
    for pre, first in (("shhd_gpu_", [P]), ("shhd_batch_", [P, ctypes.c_int])):
        f = getattr(lib, pre + "get_physical")
        f.argtypes = first + [ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
        f.restype = ctypes.c_int
        f = getattr(lib, pre + "apply_initial_condition")
        f.argtypes = first + [ctypes.POINTER(ShhdInitialCondition)]
        f.restype = ctypes.c_int
        f = getattr(lib, pre + "set_probes")
        f.argtypes = first + [ctypes.POINTER(ShhdProbe), ctypes.c_int, ctypes.c_int]
        f.restype = ctypes.c_int
        f = getattr(lib, pre + "read_probes")
        f.argtypes = first + [ctypes.POINTER(ctypes.c_double), ctypes.c_size_t, ctypes.POINTER(ctypes.c_int)]
        f.restype = ctypes.c_int
        f = getattr(lib, pre + "set_structure_factor")
        f.argtypes = first + [ctypes.c_uint, ctypes.c_int]
        f.restype = ctypes.c_int
        f = getattr(lib, pre + "get_structure_factor")
        f.argtypes = first + [ctypes.POINTER(ctypes.c_double), ctypes.c_size_t, ctypes.POINTER(ctypes.c_longlong)]
        f.restype = ctypes.c_int
    lib.shhd_batch_apply_initial_condition_noalloc.argtypes = [P, ctypes.c_int, ctypes.POINTER(ShhdInitialCondition)]
    lib.shhd_batch_apply_initial_condition_noalloc.restype = ctypes.c_int
    lib.shhd_batch_set_flux_request.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.shhd_batch_set_flux_request.restype = ctypes.c_int
    lib.shhd_batch_get_flux.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
    lib.shhd_batch_get_flux.restype = ctypes.c_int
    lib.shhd_batch_get_state.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
    lib.shhd_batch_get_state.restype = ctypes.c_int
    lib.shhd_batch_retire.argtypes = [P, ctypes.c_int]
    lib.shhd_batch_retire.restype = ctypes.c_int
    lib.shhd_batch_usage.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_size_t)]
    lib.shhd_batch_usage.restype = ctypes.c_int
    lib.shhd_batch_destroy.argtypes = [P]
    lib.shhd_batch_destroy.restype = None
    lib.shhd_batch_last_error.argtypes = [P]
    lib.shhd_batch_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


class ShhdError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _probe_array(probes):
    arr = (ShhdProbe * max(1, len(probes)))()
    for i, (src, idx, gx, gy) in enumerate(probes):
        arr[i] = ShhdProbe(int(src), int(idx), int(gx), int(gy))
    return arr


def _field_mask(fields):
    m = 0
    for f in fields:
        m |= 1 << int(f)
    return m


class _Measures:
    """Probes and structure factor (include/shhd_gpu.h, on-device FP64
    accumulators); `_call(name, *args)` adds the handle (and slot)."""

    def _set_probes(self, key, probes, capacity):
        self._probes[key] = len(probes)
        self._check(self._call("set_probes", key, _probe_array(probes), len(probes), int(capacity)))
        self._cap[key] = int(capacity)

    def _read_probes(self, key):
        n = self._probes.get(key, 0)
        out = np.zeros((self._cap.get(key, 0), n), dtype=np.complex128)
        flat = out.view(np.float64).reshape(-1)
        rows = ctypes.c_int()
        self._check(self._call("read_probes", key, _dptr(flat), flat.size, ctypes.byref(rows)))
        return out[:rows.value].copy()

    def _set_sk(self, key, fields, nevery):
        self._sk[key] = sorted(int(f) for f in fields)
        self._check(self._call("set_structure_factor", key, _field_mask(fields), int(nevery)))

    def _get_sk(self, key, cfg):
        fields = self._sk.get(key, [])
        out = np.zeros((len(fields), cfg.ny, cfg.nx // 2 + 1), dtype=np.float64)
        cnt = ctypes.c_longlong()
        flat = out.reshape(-1)
        self._check(self._call("get_structure_factor", key, _dptr(flat), flat.size, ctypes.byref(cnt)))
        return out, cnt.value


class GpuFluidSolver(_Measures):
    """Drop-in for the reference FluidSolver's step path on one HIP device."""

    KERNEL_CLASSES = ("inverse_columns", "rows", "forward_columns_stage")

    def __init__(self, cfg: SolverConfig):
        self.lib = load_library()
        self.cfg = cfg
        self._c = cfg.to_c()
        h = ctypes.c_void_p()
        st = self.lib.shhd_gpu_create(ctypes.byref(self._c), ctypes.byref(h))
        if st != 0:
            raise ShhdError(st, self.lib.shhd_gpu_last_error(None).decode())
        self.h = h
        self._probes, self._cap, self._sk = {}, {}, {}

    def _check(self, st):
        if st != 0:
            raise ShhdError(st, self.lib.shhd_gpu_last_error(self.h).decode())

    def _call(self, name, key, *args):
        return getattr(self.lib, "shhd_gpu_" + name)(self.h, *args)

    def get_physical(self) -> np.ndarray:
        """Every field on the padded physical grid, shape (N+3, My, Mx)."""
        out = np.zeros((self.cfg.num_fields, self.cfg.ny, self.cfg.nx))
        self._check(self.lib.shhd_gpu_get_physical(self.h, _dptr(out.reshape(-1)), out.size))
        return out

    def apply_initial_condition(self, ic: ShhdInitialCondition):
        """`set <target> <style>` on the device (shhd_gpu_apply_initial_condition)."""
        self._check(self.lib.shhd_gpu_apply_initial_condition(self.h, ctypes.byref(ic)))

    def set_probes(self, probes, capacity):
        """probes: list of (source, index, gx, gy); source PROBE_STATE or PROBE_FLUX."""
        self._set_probes(0, probes, capacity)

    def read_probes(self) -> np.ndarray:
        """Rows recorded since the last read: shape (rows, nprobes), complex."""
        return self._read_probes(0)

    def set_structure_factor(self, fields, nevery=1):
        self._set_sk(0, fields, nevery)

    def get_structure_factor(self):
        """(sums of |f_hat|^2, shape (nfields, My, Mx/2+1)), number of samples)."""
        return self._get_sk(0, self.cfg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.shhd_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_state(self, state: np.ndarray):
        s = np.ascontiguousarray(state, dtype=np.complex128)
        if s.shape != self.cfg.state_shape():
            raise ValueError(f"state shape {s.shape} != {self.cfg.state_shape()}")
        flat = s.view(np.float64).reshape(-1)
        self._check(self.lib.shhd_gpu_set_state(self.h, _dptr(flat), flat.size))

    def get_state(self) -> np.ndarray:
        out = np.zeros(self.cfg.state_shape(), dtype=np.complex128)
        flat = out.view(np.float64).reshape(-1)
        self._check(self.lib.shhd_gpu_get_state(self.h, _dptr(flat), flat.size))
        return out

    def step(self, nsteps: int = 1, t0: float = 0.0):
        self._check(self.lib.shhd_gpu_step(self.h, int(nsteps), float(t0)))

    def set_flux_request(self, density=False, order_parameter=False, momentum=False):
        """FluidSolver::set_flux_request (src/fluid_solver.h:101-103)."""
        self._check(self.lib.shhd_gpu_set_flux_request(self.h, int(density), int(order_parameter), int(momentum)))

    def get_flux(self) -> np.ndarray:
        """FluidSolver::flux_buffer (src/fluid_solver.h:105-107): the last step's
        fluxes, shape (2 + 2N + 3, My, Mx/2+1), slots [rho_x, rho_y,
        (psi_x, psi_y) per order parameter, Pi_xx, Pi_xy, Pi_yy]."""
        out = np.zeros(self.cfg.flux_shape(), dtype=np.complex128)
        flat = out.view(np.float64).reshape(-1)
        self._check(self.lib.shhd_gpu_get_flux(self.h, _dptr(flat), flat.size))
        return out

    def info(self) -> ShhdInfo:
        i = ShhdInfo()
        self._check(self.lib.shhd_gpu_get_info(self.h, ctypes.byref(i)))
        return i

    def set_kernel_timing(self, enable: bool):
        self._check(self.lib.shhd_gpu_set_kernel_timing(self.h, int(enable)))

    def kernel_timing(self):
        ms = (ctypes.c_double * 3)()
        n = (ctypes.c_longlong * 3)()
        self._check(self.lib.shhd_gpu_get_kernel_timing(self.h, ms, n, 3))
        return {k: (ms[i], n[i]) for i, k in enumerate(self.KERNEL_CLASSES)}

    def kernel_bytes(self):
        b = (ctypes.c_double * 3)()
        self._check(self.lib.shhd_gpu_get_kernel_bytes(self.h, b, 3))
        return {k: b[i] for i, k in enumerate(self.KERNEL_CLASSES)}


class Batch(_Measures):
    """Many resident simulations advanced together (include/shhd_gpu.h,
    shhd_batch_*).  Each admitted instance computes exactly what it would
    compute alone."""

    def __init__(self, budget_bytes: int = 0, use_graphs: bool = True):
        self.lib = load_library()
        h = ctypes.c_void_p()
        st = self.lib.shhd_batch_create(int(budget_bytes), int(bool(use_graphs)), ctypes.byref(h))
        if st != 0:
            raise ShhdError(st, self.lib.shhd_batch_last_error(None).decode())
        self.h = h
        self.cfgs = {}
        self._probes, self._cap, self._sk = {}, {}, {}

    def _check(self, st):
        if st != 0:
            raise ShhdError(st, self.lib.shhd_batch_last_error(self.h).decode())

    def _call(self, name, key, *args):
        return getattr(self.lib, "shhd_batch_" + name)(self.h, int(key), *args)

    def get_physical(self, slot) -> np.ndarray:
        cfg = self.cfgs[slot]
        out = np.zeros((cfg.num_fields, cfg.ny, cfg.nx))
        self._check(self.lib.shhd_batch_get_physical(self.h, int(slot), _dptr(out.reshape(-1)), out.size))
        return out

    def apply_initial_condition(self, slot, ic: ShhdInitialCondition):
        self._check(self.lib.shhd_batch_apply_initial_condition(self.h, int(slot), ctypes.byref(ic)))

    def apply_initial_condition_noalloc(self, slot, ic: ShhdInitialCondition):
        """The same initial condition with no device allocation on the call
        (shhd_batch_apply_initial_condition_noalloc)."""
        self._check(self.lib.shhd_batch_apply_initial_condition_noalloc(self.h, int(slot), ctypes.byref(ic)))

    def set_probes(self, slot, probes, capacity):
        self._set_probes(slot, probes, capacity)

    def read_probes(self, slot) -> np.ndarray:
        return self._read_probes(slot)

    def set_structure_factor(self, slot, fields, nevery=1):
        self._set_sk(slot, fields, nevery)

    def get_structure_factor(self, slot):
        return self._get_sk(slot, self.cfgs[slot])

    @staticmethod
    def footprint(cfg: SolverConfig) -> int:
        lib = load_library()
        c = cfg.to_c()
        n = ctypes.c_size_t()
        st = lib.shhd_batch_footprint(ctypes.byref(c), ctypes.byref(n))
        if st != 0:
            raise ShhdError(st, lib.shhd_batch_last_error(None).decode())
        return n.value

    def admit(self, cfg: SolverConfig, state: Optional[np.ndarray] = None) -> int:
        """Admit a simulation; state None starts it at zero (then apply initial conditions)."""
        c = cfg.to_c()
        slot = ctypes.c_int()
        if state is None:
            self._check(self.lib.shhd_batch_admit(self.h, ctypes.byref(c), None, 0, ctypes.byref(slot)))
        else:
            s = np.ascontiguousarray(state, dtype=np.complex128)
            if s.shape != cfg.state_shape():
                raise ValueError(f"state shape {s.shape} != {cfg.state_shape()}")
            flat = s.view(np.float64).reshape(-1)
            self._check(self.lib.shhd_batch_admit(self.h, ctypes.byref(c), _dptr(flat), flat.size,
                                                  ctypes.byref(slot)))
        self.cfgs[slot.value] = cfg
        return slot.value

    def step(self, nsteps: int):
        """Advance every admitted instance; raises ShhdError(ERR_ZERO_DENSITY)
        when an instance stopped at zero density during the call (the others
        advanced all nsteps; see failures())."""
        self._check(self.lib.shhd_batch_step(self.h, int(nsteps)))

    def hold(self, slot: int, on: bool = True):
        """Keep an instance resident without stepping it (on=False releases it)."""
        self._check(self.lib.shhd_batch_hold(self.h, int(slot), int(bool(on))))

    def failures(self):
        """{slot: 0-based failing step} of the instances stopped by zero density."""
        n = ctypes.c_int()
        self._check(self.lib.shhd_batch_failures(self.h, None, None, 0, ctypes.byref(n)))
        cap = max(n.value, 1)
        slots = (ctypes.c_int * cap)()
        steps = (ctypes.c_longlong * cap)()
        self._check(self.lib.shhd_batch_failures(self.h, slots, steps, cap, ctypes.byref(n)))
        return {int(slots[i]): int(steps[i]) for i in range(n.value)}

    def get_state(self, slot: int) -> np.ndarray:
        cfg = self.cfgs[slot]
        out = np.zeros(cfg.state_shape(), dtype=np.complex128)
        flat = out.view(np.float64).reshape(-1)
        self._check(self.lib.shhd_batch_get_state(self.h, int(slot), _dptr(flat), flat.size))
        return out

    def set_flux_request(self, slot: int, density=False, order_parameter=False, momentum=False):
        self._check(self.lib.shhd_batch_set_flux_request(self.h, int(slot), int(density), int(order_parameter),
                                                         int(momentum)))

    def get_flux(self, slot: int) -> np.ndarray:
        cfg = self.cfgs[slot]
        out = np.zeros(cfg.flux_shape(), dtype=np.complex128)
        flat = out.view(np.float64).reshape(-1)
        self._check(self.lib.shhd_batch_get_flux(self.h, int(slot), _dptr(flat), flat.size))
        return out

    def retire(self, slot: int):
        self._check(self.lib.shhd_batch_retire(self.h, int(slot)))
        self.cfgs.pop(slot, None)

    def usage(self):
        a = ctypes.c_int()
        u = ctypes.c_size_t()
        self._check(self.lib.shhd_batch_usage(self.h, ctypes.byref(a), ctypes.byref(u)))
        return a.value, u.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.shhd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------------------
# snapshots (SnapshotMeasure, src/measure_snapshot.cc)
# ---------------------------------------------------------------------------
SNAPSHOT_PHYSICAL, SNAPSHOT_SPECTRAL = 0, 1


def snapshot_filename(file: str, step: int, total_run_steps: int) -> str:
    """`<file>_step<zero-padded step>.snapshot`, width = digits of the run's
    total steps + 1 (src/measure_snapshot.cc:35-52)."""
    digits, m = 1, int(total_run_steps)
    while m >= 10:
        m //= 10
        digits += 1
    return f"{file}_step{int(step):0{digits + 1}d}.snapshot"


def write_snapshot(path: str, space: int, data: np.ndarray, cfg: SolverConfig, step: int, time: float):
    """Reference snapshot text (shhd_write_snapshot): physical data of shape
    (N+3, My, Mx) or a spectral state of shape (N+3, My, Mx/2+1)."""
    lib = load_library()
    if space == SNAPSHOT_PHYSICAL:
        arr = np.ascontiguousarray(data, dtype=np.float64)
        assert arr.shape == (cfg.num_fields, cfg.ny, cfg.nx)
        flat = arr.reshape(-1)
    else:
        arr = np.ascontiguousarray(data, dtype=np.complex128)
        assert arr.shape == cfg.state_shape()
        flat = arr.view(np.float64).reshape(-1)
    st = lib.shhd_write_snapshot(path.encode(), int(space), _dptr(flat), cfg.nx, cfg.ny, cfg.num_order_parameters,
                                 int(step), float(time))
    if st != 0:
        raise ShhdError(st, lib.shhd_gpu_last_error(None).decode())


def read_snapshot(path: str):
    """Parse a snapshot file: returns (space, step, time, data) with data shaped
    as write_snapshot takes it, so a spectral snapshot is a state again."""
    with open(path) as fh:
        head = [next(fh) for _ in range(4)]
    space = head[0].split()[-1]
    t = head[1].split()
    step, time_ = int(t[2]), float(t[4])
    d = head[2].split()
    nxs, ny, n = int(d[2]), int(d[4]), int(d[6])
    table = np.loadtxt(path, comments="#", ndmin=2)
    nf = n + 3
    if space == "physical":
        data = table[:, 2:2 + nf].T.reshape(nf, ny, nxs)
        return SNAPSHOT_PHYSICAL, step, time_, data
    vals = table[:, 2:2 + 2 * nf]
    data = (vals[:, 0::2] + 1j * vals[:, 1::2]).T.reshape(nf, ny, nxs)
    return SNAPSHOT_SPECTRAL, step, time_, data


def autotune(cfg: SolverConfig, steps: int = 6) -> dict:
    """Launch-shape autotuner (shhd_autotune): measures the candidate column
    tiles and row-pass lane counts for this problem shape and records the
    fastest for every later solver or batch instance of the shape."""
    lib = load_library()
    c = cfg.to_c()
    t = ShhdTuning()
    st = lib.shhd_autotune(ctypes.byref(c), int(steps), ctypes.byref(t))
    if st != 0:
        raise ShhdError(st, lib.shhd_gpu_last_error(None).decode())
    return {"inverse_column_tile": t.inverse_column_tile, "forward_column_tile": t.forward_column_tile,
            "row_lanes": t.row_lanes, "column_transform": t.column_transform, "row_transform": t.row_transform,
            "plane_block": t.plane_block, "inverse_plane_block": t.inverse_plane_block,
            "forward_column_waves": t.forward_column_waves, "row_waves": t.row_waves, "ms": list(t.ms)}


def set_launch_shapes(cfg: SolverConfig, inverse_column_tile: int, forward_column_tile: int, row_lanes: int,
                      column_transform: int = 0, row_transform: int = 0, plane_block: int = 0,
                      inverse_plane_block: int = 0, forward_column_waves: int = 0, row_waves: int = 0):
    """Record launch shapes measured by an earlier autotune (shhd_set_tuning)."""
    lib = load_library()
    c = cfg.to_c()
    t = ShhdTuning(int(inverse_column_tile), int(forward_column_tile), int(row_lanes), int(column_transform),
                   int(row_transform), int(plane_block), int(inverse_plane_block), int(forward_column_waves),
                   int(row_waves))
    st = lib.shhd_set_tuning(ctypes.byref(c), ctypes.byref(t))
    if st != 0:
        raise ShhdError(st, lib.shhd_gpu_last_error(None).decode())
