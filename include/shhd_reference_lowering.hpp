/*
 * shhd_reference_lowering.hpp — the reference-side half of the `gpu` backend:
 * lowers the reference solver's Params and model plugins into the device
 * library's shhd_config.  It is compiled inside the reference's source tree
 * (it includes the reference's own headers) by the FluidSolver that selects
 * the device engine; the device library itself never includes it.
 *
 * Sources of each quantity:
 *   grid, lengths, dealias        GridConfig        (src/simulationinfo.h:35-74)
 *   time_evolution, dt            RuntimeConfig     (src/simulationinfo.h:76-80),
 *                                 parse_dynamics_mode (src/fcalculator_dynamics_mode.h:13-30)
 *   cT^2                          Thermodynamics::linear_pressure_coefficient (src/model_thermodynamics.h:25-27)
 *   eta, zeta, M[a,a]             TransportCoefficient (src/model_transport_coefficient.h:11-85)
 *   mu_lin k0/k2/k4, mu_phys      FreeEnergy        (src/model_free_energy.h:27-52)
 *   nonlinear, noise, forces      FixConfig         (src/simulationinfo.h:91-124),
 *                                 noise enablement as FluidSolver (src/fluid_solver.h:79-94)
 *
 * The physical chemical potential is reachable only through
 * FreeEnergy::physical_chemical_potential(alpha, psi*) (the phi4 style keeps
 * u private, src/model_free_energy_phi4.cc:47-55).  The device evaluates it as
 * u_alpha psi_alpha^3, so the lowering reads u_alpha = mu_alpha(e_alpha) at the
 * unit vector e_alpha and accepts the style only if mu_alpha(t e_alpha + s e_beta)
 * = u_alpha t^3 at a few points (cubic in its own component, independent of
 * the others); any other physical chemical potential, physical pressure or
 * physical viscosity has no device functor and is refused.
 */
#ifndef SHHD_REFERENCE_LOWERING_HPP
#define SHHD_REFERENCE_LOWERING_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "fcalculator_dynamics_mode.h"
#include "model_free_energy.h"
#include "model_thermodynamics.h"
#include "model_transport_coefficient.h"
#include "shhd_gpu.h"
#include "simulationinfo.h"

namespace shhd_reference {

// u_alpha of a physical chemical potential of the form u_alpha psi_alpha^3,
// read through the plugin's public pointwise function; throws for any other form.
inline double phi4_cubic_coefficient(const FreeEnergy& fe, int alpha) {
    const int n = fe.num_order_parameters();
    std::vector<double> psi(static_cast<size_t>(n), 0.0);
    psi[static_cast<size_t>(alpha)] = 1.0;
    const double u = fe.physical_chemical_potential(alpha, psi.data());
    const double probes[3] = {-1.5, 0.5, 2.0};
    for (double t : probes) {
        for (int beta = 0; beta < n; ++beta) {
            std::fill(psi.begin(), psi.end(), 0.0);
            psi[static_cast<size_t>(alpha)] = t;
            if (beta != alpha) psi[static_cast<size_t>(beta)] = 0.75;
            const double mu = fe.physical_chemical_potential(alpha, psi.data());
            const double want = u * t * t * t;
            if (!(std::fabs(mu - want) <= 1e-12 * std::fmax(1.0, std::fabs(want))))
                throw std::runtime_error("gpu backend: the free energy's physical chemical potential of order parameter " +
                                         std::to_string(alpha) + " is not u psi^3 (no device functor).");
        }
    }
    return u;
}

inline shhd_config lower(const Params& p, const Thermodynamics& th, const FreeEnergy& fe,
                         const TransportCoefficient& tr) {
    if (th.has_physical_pressure())
        throw std::runtime_error("gpu backend: physical pressure has no device functor.");
    if (tr.has_physical_shear_viscosity() || tr.has_physical_bulk_viscosity())
        throw std::runtime_error("gpu backend: physical viscosity has no device functor.");
    shhd_config c{};
    c.nx_active = p.grid.active_num_grid[0];
    c.ny_active = p.grid.active_num_grid[1];
    c.nx = p.grid.num_grid[0];
    c.ny = p.grid.num_grid[1];
    c.lx = p.grid.length[0];
    c.ly = p.grid.length[1];
    const int n = p.physics.num_order_parameters;
    if (n > SHHD_MAX_ORDER_PARAMETERS) throw std::runtime_error("gpu backend: too many order parameters.");
    c.num_order_parameters = n;
    const DynamicsMode mode = parse_dynamics_mode(p.runtime.time_evolution_type);
    c.mode = is_compressible_mode(mode) ? SHHD_MODE_COMPRESSIBLE
           : is_incompressible_mode(mode) ? SHHD_MODE_INCOMPRESSIBLE : SHHD_MODE_QUIESCENT;
    c.integrator = p.runtime.time_evolution_type.rfind("srk3/", 0) == 0 ? SHHD_INTEGRATOR_SRK3 : SHHD_INTEGRATOR_EULER;
    c.dt = p.runtime.dt;
    c.precision = SHHD_PRECISION_FP64;
    c.pressure_coefficient = th.linear_pressure_coefficient();   // cT^2 (linear_eos), 0 (null)
    c.eta = tr.shear_viscosity();
    c.zeta = tr.bulk_viscosity();
    bool any_linear = false;
    for (int a = 0; a < n; ++a) {
        c.mobility[a] = tr.order_parameter_mobility()[static_cast<size_t>(a)];
        c.mu_k0[a] = fe.chemical_potential_k0_coefficient(a);
        c.mu_k2[a] = fe.chemical_potential_k2_coefficient(a);
        c.mu_k4[a] = fe.chemical_potential_k4_coefficient(a);
        any_linear = any_linear || c.mu_k0[a] != 0.0 || c.mu_k2[a] != 0.0 || c.mu_k4[a] != 0.0;
        c.phi4_u[a] = fe.has_physical_chemical_potential() ? phi4_cubic_coefficient(fe, a) : 0.0;
    }
    c.free_energy_kind = fe.has_physical_chemical_potential() ? SHHD_FREE_ENERGY_PHI4
                       : any_linear ? SHHD_FREE_ENERGY_QUADRATIC : SHHD_FREE_ENERGY_NULL;
    c.momentum_advection = p.fix.momentum_advection ? 1 : 0;
    c.order_parameter_advection = p.fix.order_parameter_advection ? 1 : 0;
    // FluidSolver wires psi_sto / j_sto only under these conditions (src/fluid_solver.h:79-94)
    const bool quiescent = is_quiescent_mode(mode);
    c.noise_order_parameter = (p.fix.noise.enabled && p.fix.noise.order_parameter_enabled && n > 0) ? 1 : 0;
    c.noise_momentum = (p.fix.noise.enabled && p.fix.noise.momentum_enabled && !quiescent) ? 1 : 0;
    c.kBT = p.fix.noise.kBT;
    c.order_parameter_noise_chi = p.fix.noise.order_parameter_noise_chi;
    c.seed = static_cast<uint64_t>(static_cast<int64_t>(p.fix.noise.seed));
    for (const SineForceFixConfig& f : p.fix.sine_forces) {
        if (!f.enabled) continue;
        if (c.num_sine_forces >= SHHD_MAX_FORCES) throw std::runtime_error("gpu backend: too many force/sine fixes.");
        c.sine_forces[c.num_sine_forces++] = {f.momentum_enabled ? 1 : 0, f.component, f.axis, f.nk, f.amplitude};
    }
    for (const GradientForceFixConfig& f : p.fix.gradient_forces) {
        if (!f.enabled) continue;
        if (c.num_gradient_forces >= SHHD_MAX_FORCES)
            throw std::runtime_error("gpu backend: too many force/gradient fixes.");
        c.gradient_forces[c.num_gradient_forces++] = {f.component, f.direction, f.amplitude};
    }
    return c;
}

}  // namespace shhd_reference

#endif
