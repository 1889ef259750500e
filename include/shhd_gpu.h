/*
 * shhd_gpu.h — C-ABI of the MI355X (gfx950) right-hand-side and time-step engine.
 *
 * The library replaces the reference solver's per-step hot path:
 *   FluidSolver::step(State&, double t)          src/fluid_solver.h:97-99
 *     -> TimeIntegrator::step(State&, t, RHSOperators)   src/time_integrator.h:141
 *        -> FCalculator::{rho_det, psi_det, j_det, psi_sto, j_sto}
 *           (src/fcalculator.h:64-69) and FCalculatorWorkspace
 *           (src/fcalculator_workspace.cc:68-145), FourierTransform2D
 *           (src/fourier_transform.h:39-51).
 *
 * Conventions shared with the reference:
 *   - State layout (src/state.cc:8-38, src/domain.cc:21-30): fields ordered
 *     [rho, psi_0 .. psi_{N-1}, jx, jy]; each field is complex[(Mx/2+1) * My],
 *     index = gy * (Mx/2+1) + gx, interleaved (re, im) doubles.
 *   - Wavenumbers (src/domain.cc:139-149): kx = 2*pi*gx/Lx,
 *     ky = 2*pi*(gy <= My/2 ? gy : gy - My)/Ly.
 *   - Active modes (src/spectral_mask.cc:45-52): gx < Nax/2 and |ky| < Nay/2.
 *   - Forward 2D transform unscaled, backward scaled by 1/(Mx*My)
 *     (src/fourier_transform.h:43,50).
 *
 * Ownership: host buffers belong to the caller; device memory belongs to the
 * handle.  A handle is bound to one HIP stream and must be used from one thread
 * at a time (the reference's FCalculator is likewise non-reentrant,
 * src/fcalculator.h:35-38).  Errors cross the ABI as status codes; the message
 * of the last failure is available from shhd_gpu_last_error().
 */
#ifndef SHHD_GPU_H
#define SHHD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHHD_MAX_ORDER_PARAMETERS 8
#define SHHD_MAX_FORCES 8

typedef enum shhd_status {
    SHHD_OK = 0,
    SHHD_ERR_INVALID = 1,      /* invalid configuration or argument            */
    SHHD_ERR_UNSUPPORTED = 2,  /* valid for the reference, no device functor   */
    SHHD_ERR_HIP = 3,          /* HIP runtime error / device unavailable       */
    SHHD_ERR_ZERO_DENSITY = 4, /* "compressible velocity encountered zero density." */
    SHHD_ERR_RUNTIME = 5,      /* other runtime error raised by the physics    */
    SHHD_ERR_FULL = 6          /* batch: device budget or instance slots used up */
} shhd_status;

/* dynamics mode: src/fcalculator_dynamics_mode.h:13-30 */
enum { SHHD_MODE_COMPRESSIBLE = 0, SHHD_MODE_INCOMPRESSIBLE = 1, SHHD_MODE_QUIESCENT = 2 };
/* integrator: src/fluid_solver.h:37-48 */
enum { SHHD_INTEGRATOR_EULER = 0, SHHD_INTEGRATOR_SRK3 = 1 };
/* free-energy style lowered to coefficients: src/model_free_energy*.cc */
enum { SHHD_FREE_ENERGY_NULL = 0, SHHD_FREE_ENERGY_QUADRATIC = 1, SHHD_FREE_ENERGY_PHI4 = 2 };
/* arithmetic of the device path */
enum { SHHD_PRECISION_FP64 = 64, SHHD_PRECISION_FP32 = 32 };

/* fix <ID> <momentum|order_parameter> force/sine on ... (src/simulationinfo.h:99-108) */
typedef struct shhd_sine_force {
    int momentum;          /* 1: momentum target, 0: order-parameter target   */
    int component;         /* momentum: 0=x 1=y; order parameter: its index   */
    int axis;              /* 0=x 1=y                                         */
    int nk;                /* 0 < nk < active_N_axis/2                        */
    double amplitude;
} shhd_sine_force;

/* fix <ID> order_parameter force/gradient on ... (src/simulationinfo.h:109-115) */
typedef struct shhd_gradient_force {
    int component;         /* order-parameter index                           */
    int direction;         /* 0=x 1=y                                         */
    double amplitude;
} shhd_gradient_force;

/*
 * Everything the hot path reads from Params and the model plugins, lowered to
 * plain numbers by the host (src/simulationinfo.h:35-173; the plugins'
 * coefficient getters, src/model_free_energy.h:27-52,
 * src/model_thermodynamics.h:25-27, src/model_transport_coefficient.h:11-85).
 */
typedef struct shhd_config {
    int nx_active, ny_active;   /* `grid` (even)                               */
    int nx, ny;                 /* compute grid after `dealias` (GridConfig)   */
    double lx, ly;              /* `length`                                    */
    int num_order_parameters;   /* N (0..SHHD_MAX_ORDER_PARAMETERS)            */
    int mode;                   /* SHHD_MODE_*                                 */
    int integrator;             /* SHHD_INTEGRATOR_*                           */
    double dt;                  /* `timestep`                                  */
    int precision;              /* SHHD_PRECISION_* (device arithmetic)        */

    double pressure_coefficient;/* cT^2 of linear_eos, 0 for null thermo       */
    double eta, zeta;           /* constant transport                          */
    double mobility[SHHD_MAX_ORDER_PARAMETERS];   /* diagonal M[a,a]          */

    int free_energy_kind;       /* SHHD_FREE_ENERGY_*                          */
    double mu_k0[SHHD_MAX_ORDER_PARAMETERS];      /* mu_lin = (k0+k2 k^2+k4 k^4) psi */
    double mu_k2[SHHD_MAX_ORDER_PARAMETERS];
    double mu_k4[SHHD_MAX_ORDER_PARAMETERS];
    double phi4_u[SHHD_MAX_ORDER_PARAMETERS];     /* mu_phys = u psi^3 (phi4)  */

    int momentum_advection;          /* fix momentum nonlinear on            */
    int order_parameter_advection;   /* fix order_parameter nonlinear on     */

    int noise_momentum;              /* fix momentum noise on                */
    int noise_order_parameter;       /* fix order_parameter noise on         */
    double kBT;                      /* default 1.0                          */
    double order_parameter_noise_chi;/* default 1.0                          */
    uint64_t seed;                   /* default 12345                        */

    int num_sine_forces;
    shhd_sine_force sine_forces[SHHD_MAX_FORCES];
    int num_gradient_forces;
    shhd_gradient_force gradient_forces[SHHD_MAX_FORCES];
} shhd_config;

/* Static description of a created handle (no reference counterpart: it
 * reports the transform counts F_in / F_out that size the byte model). */
typedef struct shhd_info {
    int num_fields;              /* N + 3                                      */
    int num_inverse_fields;      /* F_in : fields inverse-transformed per stage*/
    int num_forward_fields;      /* F_out: products forward-transformed        */
    int num_updated_fields;      /* F_s  : state fields rewritten per stage    */
    int stages_per_step;         /* 1 (Euler) or 3 (SRK3)                      */
    size_t state_doubles;        /* doubles in the reference-layout state      */
    size_t device_bytes;         /* device memory held by the handle           */
    /* launch shapes in use: kx columns per workgroup of the inverse and
     * forward column passes, lanes per row-pass workgroup, and each pass's
     * transform build: 0 staged through LDS, 1 or 2 fused (first pass from
     * HBM into registers, LDS exchanges over that many buffers) */
    int inverse_column_tile, forward_column_tile, row_lanes;
    int inverse_column_fused, forward_column_fused, row_fused;
    int inverse_column_lanes, forward_column_lanes;   /* lanes per column workgroup */
} shhd_info;

typedef struct shhd_gpu shhd_gpu;

/* FluidSolver::FluidSolver (src/fluid_solver.h:52-95): builds the integrator
 * and RHS engine.  Styles without a device functor fail here. */
shhd_status shhd_gpu_create(const shhd_config* config, shhd_gpu** out);

/* Upload a reference-layout state, (N+3)*(Mx/2+1)*My complex values given as
 * n = 2*(N+3)*(Mx/2+1)*My doubles.  Inactive modes must be zero, as every
 * reference initial condition writes active modes only
 * (src/initial_condition_*.cc loops over SpectralMask2D::active_modes()). */
shhd_status shhd_gpu_set_state(shhd_gpu* h, const double* interleaved, size_t n);

/* Download the state in the reference layout (inactive modes written as 0). */
shhd_status shhd_gpu_get_state(shhd_gpu* h, double* out, size_t n);

/* nsteps calls of FluidSolver::step(state, t0 + i*dt) (src/project.h:128-145,
 * src/fluid_solver.h:97-99).  Returns after the device work has finished.
 * When the density reaches zero (src/fcalculator_workspace.cc:126-128) the
 * call stops at that step and returns SHHD_ERR_ZERO_DENSITY with the state as
 * of the failing step's start, as the reference's throw leaves it. */
shhd_status shhd_gpu_step(shhd_gpu* h, int nsteps, double t0);

/* FluidSolver::set_flux_request (src/fluid_solver.h:101-103, FluxRequest of
 * src/buffer_flux.h).  While any flag is set, every step records the
 * step-averaged spectral fluxes of the requested conservation laws
 * (FluxBuffer::accumulate_stage_flux with stage weights 1/6, 1/6, 2/3 for SRK3
 * or 1 for Euler, draw A of the noise flux scaled by 1/sqrt(dt), and the
 * incompressible projection pressure on Pi_xx and Pi_yy). */
shhd_status shhd_gpu_set_flux_request(shhd_gpu* h, int density, int order_parameter, int momentum);
/* FluidSolver::flux_buffer (src/fluid_solver.h:105-107): the flux of the last
 * step as 2 + 2N + 3 slots [rho_x, rho_y, (psi_x, psi_y) per order parameter,
 * Pi_xx, Pi_xy, Pi_yy] (src/buffer_flux.cc:10-48), each complex[(Mx/2+1)*My]
 * in the state's layout, n = 2 * (2N+5) * (Mx/2+1) * My doubles.  Slots that
 * are not requested read 0. */
shhd_status shhd_gpu_get_flux(shhd_gpu* h, double* out, size_t n);

/* ---- on-device FP64 accumulators -------------------------------------------
 * Measures that read a few modes every step (TimeSeriesMeasure's flux columns,
 * src/measure_time_series.cc:189-227; YokotaGreenKuboMeasure's Pi_xy at chosen
 * wave vectors, src/measure_yokota_green_kubo.cc:128-193) or a per-mode average
 * (structure factor) are fed on the device, with no per-step download. */
typedef struct shhd_probe {
    int source;            /* 0: state field, 1: flux slot (see shhd_gpu_get_flux) */
    int index;             /* field or slot index                              */
    int gx, gy;            /* reference-layout mode, inside the active mask    */
} shhd_probe;

/* After every step, append the values of the probed modes (as FP64 complex)
 * to a ring of `capacity` rows.  nprobes = 0 removes the probes.  A step that
 * would overflow the ring fails with SHHD_ERR_FULL before doing any work. */
shhd_status shhd_gpu_set_probes(shhd_gpu* h, const shhd_probe* probes, int nprobes, int capacity);
/* Copy the rows recorded since the last read, oldest first, into out
 * (rows * nprobes * 2 doubles, n >= that) and empty the ring. */
shhd_status shhd_gpu_read_probes(shhd_gpu* h, double* out, size_t n, int* nrows);
/* Every `nevery` steps from now on add |f_hat(k)|^2 of each field in
 * field_mask (bit f = state field f) to an FP64 sum per active mode.
 * field_mask = 0 removes it; a new call restarts the sums. */
shhd_status shhd_gpu_set_structure_factor(shhd_gpu* h, unsigned field_mask, int nevery);
/* The sums, one real[(Mx/2+1) * My] per selected field in field order
 * (reference layout, inactive modes 0), and the number of samples. */
shhd_status shhd_gpu_get_structure_factor(shhd_gpu* h, double* out, size_t n, long long* nsamples);

/* ---- device-side initial conditions -----------------------------------------
 * `set <target> <style> ...` (src/initial_condition_*.cc) evaluated on the
 * device straight into the resident state: the mode-hash styles are pure
 * functions of (seed, channel, gx, ky, draw) through splitmix64, so a large
 * grid needs no host generation and no upload.  Like the reference's apply(),
 * each call overwrites one field (every mode; only active modes nonzero). */
enum {
    /* set density equilibrium/gaussian rho0 kBT cT seed
       (src/initial_condition_equilibrium_gaussian_density.cc:56-93) */
    SHHD_IC_EQUILIBRIUM_DENSITY = 0,
    /* set order_parameter <a> equilibrium/gaussian psi0 kBT k0 k2 k4 seed
       (src/initial_condition_equilibrium_gaussian_order_parameter.cc:59-108) */
    SHHD_IC_EQUILIBRIUM_ORDER_PARAMETER = 1,
    /* set momentum <x|y> equilibrium/gaussian/compressible kBT seed
       (src/initial_condition_equilibrium_gaussian_momentum_compressible.cc:76-112) */
    SHHD_IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE = 2,
    /* set momentum <x|y> equilibrium/gaussian/incompressible kBT seed
       (src/initial_condition_equilibrium_gaussian_momentum_incompressible.cc:76-115) */
    SHHD_IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE = 3,
    /* set momentum <x|y> random_vorticity n0 sigma urms seed
       (src/initial_condition_random_vorticity_momentum.cc:101-147) */
    SHHD_IC_RANDOM_VORTICITY = 4,
    /* No reference counterpart: transverse random-phase velocity
       u_hat = i (-ky, kx)/|k| n^(3/2) exp(-(n/n0)^2) e^{i theta} over integer
       wave numbers n = |(gx, ky)| (energy spectrum E(n) ∝ n^4 exp(-2 (n/n0)^2)),
       phases hashed as random_vorticity's, normalized like it to
       <|u|^2> = urms^2, times rho0 (the incompressible-turbulence start
       E(k) ∝ k^4 exp(-2(k/8)^2)). */
    SHHD_IC_SPECTRUM_K4 = 5
};

typedef struct shhd_initial_condition {
    int kind;                 /* SHHD_IC_*                                        */
    int component;            /* order-parameter index, or momentum direction 0/1  */
    int seed;
    double kBT;               /* equilibrium styles                              */
    double value;             /* rho0 (density) or psi0 (order parameter)        */
    double pressure_coefficient; /* density: cT^2                                */
    double k0, k2, k4;        /* order parameter: linear chemical-potential coefficients */
    double n0, sigma, urms;   /* random_vorticity (n0, sigma, urms); spectrum (n0, urms) */
} shhd_initial_condition;

/* Momentum styles read rho0 = Re rho_hat(0,0) / (Mx*My) from the resident
 * state, as the reference does, so density is set first.  Errors carry the
 * reference's messages. */
shhd_status shhd_gpu_apply_initial_condition(shhd_gpu* h, const shhd_initial_condition* ic);

/* ---- snapshots (SnapshotMeasure, src/measure_snapshot.cc) -------------------- */

/* Every field in physical space on the padded grid, as the reference's
 * PhysicalStateBuffer after FourierTransform2D::backward_many
 * (src/measure_workspace.h:27-35, src/buffer_physical_state.h): out[f][gy][gx],
 * n = (N+3) * Mx * My doubles, computed on the device (inverse column pass of
 * every field, then the row half). */
shhd_status shhd_gpu_get_physical(shhd_gpu* h, double* out, size_t n);

enum { SHHD_SNAPSHOT_PHYSICAL = 0, SHHD_SNAPSHOT_SPECTRAL = 1 };
/* Write `data` as the reference's snapshot text (write_physical_snapshot /
 * write_spectral_snapshot, src/measure_snapshot.cc:54-138): physical data as
 * from shhd_gpu_get_physical, spectral data as from shhd_gpu_get_state;
 * nx, ny = the padded grid Mx, My.  Host-only: touches no device. */
shhd_status shhd_write_snapshot(const char* path, int space, const double* data, int nx, int ny,
                                int num_order_parameters, int step, double time);

/* Destroy the handle and free its device memory. */
void shhd_gpu_destroy(shhd_gpu* h);

/* Message of the last failing call on h (or of the last failing create when
 * h is NULL). */
const char* shhd_gpu_last_error(const shhd_gpu* h);

/* ---- instrumentation (no reference counterpart) ---------------------------- */

shhd_status shhd_gpu_get_info(const shhd_gpu* h, shhd_info* info);

/* Per-kernel timing with HIP events recorded on the handle's stream.  While
 * enabled, every kernel launch of shhd_gpu_step is bracketed by events; the
 * accumulated time per kernel class (0: inverse column pass, 1: row pass,
 * 2: forward column pass + stage update) is returned in ms together with the
 * launch counts.  Enabling resets the accumulators. */
shhd_status shhd_gpu_set_kernel_timing(shhd_gpu* h, int enable);
shhd_status shhd_gpu_get_kernel_timing(shhd_gpu* h, double* total_ms, long long* launches, int nclasses);

/* ---- launch-shape autotuner (no reference counterpart) ------------------------
 * Times the candidate column tiles of the inverse and forward column passes,
 * the candidate lane counts of the row pass and the transform builds / radix
 * orders of the column and row FFTs for this configuration's
 * problem shape on a scratch instance (`steps` timed steps per candidate),
 * and records the fastest of each for the process: every handle or batch
 * instance of that shape created afterwards uses them.  Results are the same
 * whatever the shapes (only the workgroup decomposition changes). */
typedef struct shhd_tuning {
    int inverse_column_tile;   /* kx columns per inverse-column workgroup     */
    int forward_column_tile;   /* kx columns per forward-column workgroup     */
    int row_lanes;             /* lanes per row-pass workgroup                */
    int column_transform;      /* column FFTs: 0 fixed-length build staged
                                  through LDS where one exists (else the
                                  runtime plan), 1 runtime plan radix 8 first,
                                  2 runtime plan radix 4, 3 and 4 fused
                                  fixed-length build with one and two LDS
                                  buffers (where one is built for the length
                                  and tile, else 0)                           */
    int row_transform;         /* row FFTs, the same choices                  */
    int plane_block;           /* log2 of the kx columns per block of the
                                  row-to-column planes (0: column-major; a
                                  block that does not divide Nax/2 uses 0).
                                  A shape never tuned or set runs the fused
                                  builds with two LDS buffers where two
                                  workgroups still fit a CU (4, 4) and blocks
                                  of four columns (2).                       */
    int inverse_plane_block;   /* kx blocks of the column-to-row planes (the
                                  inverse column pass writes them, the row
                                  pass reads them): 0 row-major (the default),
                                  k > 0 blocks of 2^(k-1) columns laid out as
                                  the row-to-column planes' (a block that
                                  does not divide Nax/2 uses row-major)      */
    int forward_column_waves;  /* forward-column workgroups of one-wave fused
                                  tiles: 0 four wave groups splitting the
                                  products where a launch has few workgroups
                                  (the default), 1 never, 4 wherever built */
    int row_waves;             /* row-pass workgroups of one-wave staged row
                                  tiles: the same choices                    */
    double ms[3];              /* mean launch time of each kernel class (as
                                  shhd_gpu_get_kernel_timing) with that shape */
} shhd_tuning;
shhd_status shhd_autotune(const shhd_config* config, int steps, shhd_tuning* out);
/* Record launch shapes measured earlier (ms ignored; 0 entries: default). */
shhd_status shhd_set_tuning(const shhd_config* config, const shhd_tuning* tuning);

/* Algorithmic HBM bytes moved by one launch of each kernel class (the byte
 * model: compact spectra s*Na^2 per field, column/row intermediates
 * 1.5*s*Na^2 per field under 3/2 padding). */
shhd_status shhd_gpu_get_kernel_bytes(const shhd_gpu* h, double* bytes, int nclasses);

/* ---- batched instances ------------------------------------------------------
 * No single reference counterpart: the reference runs ensembles as independent
 * one-rank processes (examples/05_yokota_green_kubo/job_ohtaka_linear_samples.sh:26-36,
 * dev_env/ohtaka/README.md:283-293).  A batch keeps many simulations resident
 * in HBM and advances all of them with one launch per kernel and stage for each
 * launch shape; every instance computes exactly what it would compute alone
 * (same kernels, per-instance Philox counters), whatever else is in the batch.
 * Instances may differ in grid, model, integrator and precision.
 */
typedef struct shhd_batch shhd_batch;

/* device_budget_bytes = 0: no budget.  use_graphs: replay captured 3-step
 * hipGraphs for runs of >= 3 steps. */
shhd_status shhd_batch_create(size_t device_budget_bytes, int use_graphs, shhd_batch** out);
/* device bytes an instance of this configuration would hold */
shhd_status shhd_batch_footprint(const shhd_config* config, size_t* bytes);
/* admit one simulation with its initial reference-layout state (NULL: every
 * field zero, as a new reference State, for shhd_batch_apply_initial_condition);
 * SHHD_ERR_FULL when the budget or the slots are used up */
shhd_status shhd_batch_admit(shhd_batch* b, const shhd_config* config, const double* state, size_t n, int* slot);
/* advance every admitted instance by nsteps.  An instance whose density
 * reaches zero stops at the failing step's start (as shhd_gpu_step) and takes
 * no further step until retired; the others advance all nsteps.  The call
 * returns SHHD_ERR_ZERO_DENSITY when an instance stopped during it. */
shhd_status shhd_batch_step(shhd_batch* b, int nsteps);
/* hold = 1: the instance keeps its state and takes no step in later
 * shhd_batch_step calls until released with hold = 0 (a scheduler runs the
 * jobs it chooses and leaves the others resident) */
shhd_status shhd_batch_hold(shhd_batch* b, int slot, int hold);
/* instances stopped by zero density: up to `capacity` slots and, for each,
 * the 0-based index of the step that failed (= steps completed); *count is
 * the total number */
shhd_status shhd_batch_failures(const shhd_batch* b, int* slots, long long* steps, int capacity, int* count);
shhd_status shhd_batch_get_state(shhd_batch* b, int slot, double* out, size_t n);
/* flux recording of one instance, as shhd_gpu_set_flux_request / shhd_gpu_get_flux */
shhd_status shhd_batch_set_flux_request(shhd_batch* b, int slot, int density, int order_parameter, int momentum);
shhd_status shhd_batch_get_flux(shhd_batch* b, int slot, double* out, size_t n);
/* on-device accumulators of one instance, as the shhd_gpu_* versions */
shhd_status shhd_batch_set_probes(shhd_batch* b, int slot, const shhd_probe* probes, int nprobes, int capacity);
shhd_status shhd_batch_read_probes(shhd_batch* b, int slot, double* out, size_t n, int* nrows);
shhd_status shhd_batch_set_structure_factor(shhd_batch* b, int slot, unsigned field_mask, int nevery);
shhd_status shhd_batch_get_structure_factor(shhd_batch* b, int slot, double* out, size_t n, long long* nsamples);
/* initial condition of one instance, as shhd_gpu_apply_initial_condition */
shhd_status shhd_batch_apply_initial_condition(shhd_batch* b, int slot, const shhd_initial_condition* ic);
/* No counterpart in the reference: the same initial condition as
 * shhd_batch_apply_initial_condition, for callers that cannot allocate device
 * memory (a serving loop seeding admitted jobs between graph replays, code
 * under allocation bans).  It calls no hipMalloc or hipFree: the velocity
 * normalization of the random_vorticity and spectrum/k4 styles keeps its
 * partial sums (one per workgroup, at most 2048 FP64 values, 16 KiB) in a
 * static device array of the library that every device keeps until the
 * process ends.  Calls that overlap on one device race on that array. */
shhd_status shhd_batch_apply_initial_condition_noalloc(shhd_batch* b, int slot, const shhd_initial_condition* ic);
/* physical fields of one instance, as shhd_gpu_get_physical */
shhd_status shhd_batch_get_physical(shhd_batch* b, int slot, double* out, size_t n);
/* free an instance's device memory and slot */
shhd_status shhd_batch_retire(shhd_batch* b, int slot);
shhd_status shhd_batch_usage(const shhd_batch* b, int* active, size_t* used_bytes);
void        shhd_batch_destroy(shhd_batch* b);
const char* shhd_batch_last_error(const shhd_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* SHHD_GPU_H */
