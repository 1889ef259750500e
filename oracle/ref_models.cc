// ref_models.cc — TEST INFRASTRUCTURE ONLY.  Compiled, with the reference's
// own model-plugin sources taken where they lie (oracle/Makefile.ref), into
// oracle/_ref/libref_models.so, and loaded only by tests/: it lets the CPU
// tests run the reference's Phi4/Quadratic/Null free energies, linear-EOS and
// null thermodynamics and constant transport, and the reference-side lowering
// of include/shhd_reference_lowering.hpp on them.  Nothing in the product path
// loads it.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "model_free_energy_null.h"
#include "model_free_energy_phi4.h"
#include "model_free_energy_quadratic.h"
#include "model_thermodynamics_linear_eos.h"
#include "model_thermodynamics_null.h"
#include "model_transport_coefficient_constant.h"
#include "shhd_reference_lowering.hpp"

extern "C" {

// What an input script sets on the path (src/param_parser.cc command
// semantics), in plain numbers.
typedef struct ref_model_spec {
    int nx_active, ny_active;
    int dealias;                 // 0 none, 1 three_halves, 2 two
    double lx, ly;
    int num_order_parameters;
    char time_evolution[32];     // e.g. "srk3/compressible"
    double dt;
    int thermo;                  // 0 null, 1 linear_eos
    double cT;
    int free_energy;             // 0 null, 1 quadratic (a), 2 phi4 (a, b, u)
    double a[SHHD_MAX_ORDER_PARAMETERS], b[SHHD_MAX_ORDER_PARAMETERS], u[SHHD_MAX_ORDER_PARAMETERS];
    double eta, zeta;
    double mobility[SHHD_MAX_ORDER_PARAMETERS];
    int momentum_advection, order_parameter_advection;
    int noise_momentum, noise_order_parameter;   // `fix ... noise on` targets
    int seed;
    double kBT, chi;
} ref_model_spec;

}

namespace {

std::string g_error;

std::unique_ptr<FreeEnergy> make_free_energy(const ref_model_spec& s) {
    const int n = s.num_order_parameters;
    std::vector<double> a(s.a, s.a + n), b(s.b, s.b + n), u(s.u, s.u + n);
    if (s.free_energy == 2) return std::make_unique<Phi4FreeEnergy>(n, a, b, u);
    if (s.free_energy == 1) return std::make_unique<QuadraticFreeEnergy>(n, a);
    return std::make_unique<NullFreeEnergy>(n);
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_error.c_str(); }

// Params and plugin objects as the reference builds them, lowered by the
// reference-side lowering; 0 on success, 1 when it refuses (ref_last_error).
int ref_lower(const ref_model_spec* s, shhd_config* out) {
    try {
        Params p;
        p.grid.active_num_grid[0] = s->nx_active;
        p.grid.active_num_grid[1] = s->ny_active;
        p.grid.length[0] = s->lx;
        p.grid.length[1] = s->ly;
        p.grid.dealias_rule = s->dealias == 1 ? DealiasRule::ThreeHalves
                            : s->dealias == 2 ? DealiasRule::Two : DealiasRule::None;
        p.grid.update_compute_grid();
        p.physics.num_order_parameters = s->num_order_parameters;
        p.runtime.time_evolution_type = s->time_evolution;
        p.runtime.dt = s->dt;
        p.fix.momentum_advection = s->momentum_advection != 0;
        p.fix.order_parameter_advection = s->order_parameter_advection != 0;
        // `fix <id> <target> noise on ...` (src/param_parser.cc:625-638)
        const bool quiescent = parse_dynamics_mode(p.runtime.time_evolution_type) == DynamicsMode::Quiescent;
        p.fix.noise.momentum_enabled = s->noise_momentum != 0 && !quiescent;
        p.fix.noise.order_parameter_enabled = s->noise_order_parameter != 0 && s->num_order_parameters > 0;
        p.fix.noise.enabled = p.fix.noise.momentum_enabled || p.fix.noise.order_parameter_enabled;
        p.fix.noise.seed = s->seed;
        p.fix.noise.kBT = s->kBT;
        p.fix.noise.order_parameter_noise_chi = s->chi;
        std::unique_ptr<Thermodynamics> th;
        if (s->thermo == 1) th = std::make_unique<LinearEOSThermodynamics>(s->cT);
        else th = std::make_unique<NullThermodynamics>();
        const std::unique_ptr<FreeEnergy> fe = make_free_energy(*s);
        const ConstantTransportCoefficient tr(s->num_order_parameters, s->eta, s->zeta,
                                              std::vector<double>(s->mobility, s->mobility + s->num_order_parameters));
        *out = shhd_reference::lower(p, *th, *fe, tr);
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    }
}

// The reference free energy's pointwise physical chemical potential of order
// parameter alpha at npts points; psi holds the N fields point-major
// (psi[i * N + beta]).
int ref_physical_chemical_potential(const ref_model_spec* s, int alpha, const double* psi, size_t npts, double* out) {
    try {
        const std::unique_ptr<FreeEnergy> fe = make_free_energy(*s);
        const int n = s->num_order_parameters;
        for (size_t i = 0; i < npts; ++i) out[i] = fe->physical_chemical_potential(alpha, psi + i * static_cast<size_t>(n));
        return 0;
    } catch (const std::exception& e) {
        g_error = e.what();
        return 1;
    }
}

}
