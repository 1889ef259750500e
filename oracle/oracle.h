/*
 * oracle.h — C entry points of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle restates, on the CPU in FP64, the
 * reference solver's right-hand-side and time-step path (FCalculator,
 * FCalculatorWorkspace, FourierTransform2D, SpectralMask2D and the six
 * TimeIntegrator subclasses).  It is the checker of the device library: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Nothing in the product path links or calls it.
 */
#ifndef SHHD_ORACLE_H
#define SHHD_ORACLE_H

#include <stddef.h>
#include "../include/shhd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_solver oracle_solver;

int  oracle_create(const shhd_config* cfg, oracle_solver** out);
const char* oracle_create_error(void);
void oracle_destroy(oracle_solver* s);
const char* oracle_last_error(const oracle_solver* s);

int  oracle_set_state(oracle_solver* s, const double* interleaved, size_t n);
int  oracle_get_state(const oracle_solver* s, double* out, size_t n);
int  oracle_step(oracle_solver* s, int nsteps, double t0);

/* Deterministic RHS of every field for the given state, exactly as one
 * integrator stage would evaluate it (rho_det, psi_det for each order
 * parameter, j_det).  Fields that the mode does not evaluate are zero. */
int  oracle_rhs_det(oracle_solver* s, const double* state, double* out, double t);
/* One stochastic draw: psi_sto for each order parameter, then j_sto, from the
 * solver's mt19937_64 stream. */
int  oracle_rhs_sto(oracle_solver* s, double* out);

/* FluxBuffer (src/buffer_flux.{h,cc}): request and read the step-averaged
 * spectral fluxes, slots [rho x, rho y, (psi x, psi y)*N, Pi xx, Pi xy, Pi yy]. */
int  oracle_set_flux_request(oracle_solver* s, int density, int order_parameter, int momentum);
int  oracle_get_flux(const oracle_solver* s, double* out, size_t n);

/* `set` initial conditions (src/initial_condition_*.cc) applied to the
 * oracle's state; splitmix64 as the mode hash uses it. */
int  oracle_apply_initial_condition(oracle_solver* s, const shhd_initial_condition* ic);
unsigned long long oracle_splitmix64(unsigned long long v);

/* Transforms used by the oracle (pinned against the DFT definition). */
int  oracle_fft_c2c(int n, const double* in, double* out, int sign);
int  oracle_fft2_r2c(int mx, int my, const double* phys, double* spec);
int  oracle_fft2_c2r(int mx, int my, const double* spec, double* phys);

/* OpenMP threads used by the oracle (1 = scalar); returns the count in use. */
int  oracle_set_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
