"""Fused fixed-length transforms (transform choices 3 and 4) compute what the
staged fixed-length builds compute.

In a fused build the first Stockham pass of every transform loads its inputs
from HBM straight into registers, the row pass hands the radix-3 outputs of
its inverse transforms to the radix-3 first pass of its forward transforms in
registers, the forward column pass accumulates the radix-3 outputs of its
column transforms in registers, and the passes in between exchange through
one or two LDS buffers.  Only the order of the floating-point operations
inside a transform changes, so the states after a few SRK3 steps (noise,
flux recording and the Hermitian projection on) agree to rounding: 1e-12
relative L2 in FP64, 1e-5 in the FP32 form.  The staged builds themselves
are checked against the oracle by the parity tests.
"""
import dataclasses

import numpy as np
import pytest

import shhd
from helpers import flux_state, rel_l2

pytestmark = pytest.mark.gpu


def model_h(nx, ny, fp32=False, mode=None, dealias="3/2"):
    mode = shhd.MODE_COMPRESSIBLE if mode is None else mode
    extra = {} if mode == shhd.MODE_QUIESCENT else dict(pressure_coefficient=1.0, eta=0.5, zeta=0.3,
                                                        momentum_advection=True, noise_momentum=True)
    if mode == shhd.MODE_INCOMPRESSIBLE:
        extra.pop("pressure_coefficient")
    return shhd.SolverConfig(nx_active=nx, ny_active=ny, dealias=dealias, lx=float(nx), ly=1.1 * ny,
                             num_order_parameters=2, mobility=[1.0, 1.3], free_energy_kind=shhd.FREE_ENERGY_PHI4,
                             mu_k0=[-1.0, -1.0], mu_k2=[1.0, 1.0], phi4_u=[1.0, 1.0], mode=mode,
                             order_parameter_advection=mode != shhd.MODE_QUIESCENT, noise_order_parameter=True,
                             kBT=1e-4, dt=0.005, precision=32 if fp32 else 64, **extra)


def run(cfg, st, shapes, steps=3):
    shhd.set_launch_shapes(cfg, *shapes)
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_flux_request(density=True, order_parameter=True, momentum=True)
        g.set_state(st)
        g.step(steps)
        info = g.info()
        return g.get_state(), g.get_flux(), info


def compare(cfg, st, ref, out, tol):
    for f in range(cfg.num_fields):
        assert rel_l2(out[0][f], ref[0][f]) <= tol, f
    for s in range(ref[1].shape[0]):
        assert rel_l2(out[1][s], ref[1][s]) <= tol, ("flux", s)


# (grid, column tiles with a fused build at that length, fused row build?)
CASES = [
    ((128, 128), (2, 8), False),    # columns of 192
    ((256, 128), (2, 8), False),    # rows of 384, columns of 192
    ((512, 256), (1, 4), True),     # rows of 768, columns of 384
    ((1024, 512), (2, 4), True),    # rows of 1536, columns of 768
]


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("grid,tiles,rows_fused", CASES)
def test_fused_transforms_match_staged(grid, tiles, rows_fused, fp32):
    print(f"[case] fused transforms {grid} fp32={fp32}", flush=True)
    cfg = model_h(*grid, fp32=fp32)
    st = flux_state(cfg)
    tol = 1e-5 if fp32 else 1e-12
    ref = run(cfg, st, (tiles[0], tiles[0], 0, 0, 0))
    assert (ref[2].inverse_column_fused, ref[2].forward_column_fused, ref[2].row_fused) == (0, 0, 0)
    for c in tiles:
        for ct, rt in ((3, 4), (4, 3), (4, 4)):
            out = run(cfg, st, (c, c, 0, ct, rt))
            info = out[2]
            assert info.inverse_column_tile == c and info.forward_column_tile == c
            # two buffers where two workgroups still fit a CU's LDS, else one
            assert 1 <= info.inverse_column_fused <= ct - 2 and 1 <= info.forward_column_fused <= ct - 2
            assert (1 <= info.row_fused <= rt - 2) if rows_fused else info.row_fused == 0
            compare(cfg, st, ref, out, tol)


@pytest.mark.parametrize("integrator", [shhd.INTEGRATOR_SRK3, shhd.INTEGRATOR_EULER])
@pytest.mark.parametrize("mode", [shhd.MODE_COMPRESSIBLE, shhd.MODE_INCOMPRESSIBLE, shhd.MODE_QUIESCENT])
def test_fused_other_modes(mode, integrator):
    print(f"[case] fused transforms mode {mode} integrator {integrator}", flush=True)
    cfg = dataclasses.replace(model_h(512, 512, mode=mode), integrator=integrator)
    st = flux_state(cfg)
    ref = run(cfg, st, (2, 2, 0, 0, 0))
    out = run(cfg, st, (2, 2, 0, 4, 4))
    assert out[2].forward_column_fused == 2 and out[2].row_fused == 2
    compare(cfg, st, ref, out, 1e-12)


def test_fused_needs_the_padded_thirds():
    """Without 3/2 padding the kept ky reach into the middle third of a
    column, which the fused forward column build does not keep: the planner
    stays with the staged build."""
    print("[case] fused transforms dealias none", flush=True)
    cfg = model_h(192, 192, dealias="none")
    st = flux_state(cfg)
    ref = run(cfg, st, (1, 1, 0, 0, 0))
    out = run(cfg, st, (1, 1, 0, 4, 4))
    assert out[2].forward_column_fused == 0
    compare(cfg, st, ref, out, 1e-12)


@pytest.mark.parametrize("fp32", [False, True])
def test_plane_blocks_give_identical_results(fp32):
    """The kx block of the row-to-column planes only moves where the row pass
    stores a value and the forward column pass finds it: the states are
    bit-identical for every block, staged and fused."""
    print(f"[case] plane blocks fp32={fp32}", flush=True)
    cfg = model_h(512, 256, fp32=fp32)
    st = flux_state(cfg)
    for ct, rt in ((0, 0), (4, 4)):
        ref = run(cfg, st, (1, 1, 0, ct, rt, 0))
        for pb in (1, 2, 3):
            out = run(cfg, st, (1, 1, 0, ct, rt, pb))
            for f in range(cfg.num_fields):
                assert np.array_equal(out[0][f], ref[0][f]), (ct, pb, f)
            assert np.array_equal(out[1], ref[1]), (ct, pb)


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("grid", [(128, 128), (512, 256)])
def test_inverse_plane_blocks_give_identical_results(fp32, grid):
    """The kx block of the column-to-row planes only moves where the inverse
    column pass stores a value and the row pass (and the physical-field row
    pass) finds it: states, fluxes and physical fields are bit-identical for
    row-major planes and every block, staged and fused (rows of 192: staged
    row pass; rows of 768: fused)."""
    print(f"[case] inverse plane blocks fp32={fp32} grid={grid}", flush=True)
    cfg = model_h(*grid, fp32=fp32)
    st = flux_state(cfg)

    def run_phys(shapes):
        shhd.set_launch_shapes(cfg, *shapes)
        with shhd.GpuFluidSolver(cfg) as g:
            g.set_flux_request(density=True, order_parameter=True, momentum=True)
            g.set_state(st)
            g.step(3)
            return g.get_state(), g.get_flux(), g.get_physical()

    for ct, rt in ((0, 0), (4, 4)):
        ref = run_phys((2, 1, 0, ct, rt, 1, 0))
        for tb in (1, 2, 3, 4):
            out = run_phys((2, 1, 0, ct, rt, 1, tb))
            for k in range(3):
                assert np.array_equal(out[k], ref[k]), (ct, tb, k)



@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("na", [64, 128, 256])
def test_wave_groups_match_one_wave_tiles(fp32, na):
    """Short columns and rows on few workgroups split their transforms over
    four wave groups (forward_column_waves / row_waves = 4).  The row pass
    then computes exactly what one wave computes (bit for bit); the forward
    column pass adds the groups' partial sums of each mode, so the states
    agree to rounding, and with the step flux recorded (group 0 takes every
    product) bit for bit, fluxes included."""
    print(f"[case] wave groups na={na} fp32={fp32}", flush=True)
    cfg = model_h(na, na, fp32=fp32)
    st = flux_state(cfg)
    tol = 1e-5 if fp32 else 1e-12

    def go(fw, rw, flux):
        shhd.set_launch_shapes(cfg, 0, 0, 0, 3, 3, 1, 0, fw, rw)
        with shhd.GpuFluidSolver(cfg) as g:
            if flux:
                g.set_flux_request(density=True, order_parameter=True, momentum=True)
            g.set_state(st)
            g.step(4)
            return g.get_state(), (g.get_flux() if flux else None), g.info()

    for flux in (False, True):
        ref = go(1, 1, flux)
        assert ref[2].row_lanes == 64   # one-wave row tiles at these lengths
        # the forward column pass splits fused one-wave tiles only (192-point
        # columns in both precisions; 384-point FP64 columns)
        split3 = ref[2].forward_column_fused > 0 and ref[2].forward_column_lanes == 64
        assert split3 or na != 128
        for fw, rw in ((4, 1), (1, 4), (4, 4)):
            out = go(fw, rw, flux)
            assert out[2].forward_column_lanes == (256 if fw == 4 and split3 else ref[2].forward_column_lanes)
            assert out[2].row_lanes == (256 if rw == 4 else 64)
            if flux or fw == 1:
                assert np.array_equal(out[0], ref[0]), (fw, rw, flux)
                if flux:
                    assert np.array_equal(out[1], ref[1]), (fw, rw)
            else:
                for f in range(cfg.num_fields):
                    assert rel_l2(out[0][f], ref[0][f]) <= tol, (fw, rw, f)
