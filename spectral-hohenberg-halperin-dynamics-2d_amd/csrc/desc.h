// desc.h — the per-instance descriptor that the host lowers from shhd_config,
// and the device-side tables through which one launch serves many instances.
#pragma once

#include "lds_fft.h"

namespace shhd {

constexpr int kMaxOp = 8;                // SHHD_MAX_ORDER_PARAMETERS
constexpr int kMaxFields = kMaxOp + 3;   // rho, psi_0..psi_{N-1}, jx, jy
constexpr int kMaxIn = kMaxOp + 3;       // inverse-transformed fields
constexpr int kMaxOut = 4 * kMaxOp + 5;  // forward-transformed products
constexpr int kMaxGroups = kMaxOp + 1;   // one per psi_alpha + momentum
constexpr int kMaxContrib = 5;
constexpr int kMaxSources = 16;

// Roles of the physical fields held per point in the row pass.
enum InRole : int { ROLE_RHO = 0, ROLE_JX = 1, ROLE_JY = 2, ROLE_PSI0 = 3 };

// Products formed in the row pass and forward transformed (the reference's
// temporary_physical_fields contents, src/fcalculator.cc:189-494, 776-840).
enum OutKind : int {
    OUT_MU = 0,      // u_a psi_a^3                        (phi4 mu_phys)
    OUT_PSIVX = 1,   // psi_a v_x
    OUT_PSIVY = 2,   // psi_a v_y
    OUT_GRADF = 3,   // sum_f -A_f v_dir(f)   (compressible force/gradient)
    OUT_VX = 4,      // v_x                   (compressible viscous term)
    OUT_VY = 5,
    OUT_PXX = 6,     // j_x v_x               (momentum advection)
    OUT_PXY = 7,     // j_x v_y
    OUT_PYY = 8      // j_y v_y
};

enum GroupKind : int { GROUP_PSI = 0, GROUP_MOM = 1 };

struct OutItem {
    int kind;
    int alpha;
};

// What one forward-transformed product v_hat adds at mode k, written as
// polynomials in (kx, ky) so the forward-column kernel evaluates every kind of
// product with the same few multiply-adds:
//   acc_r += (a_r0 + a_r1 kx^2 + a_r2 ky^2 + a_r3 kx ky) v - i (b_r0 kx + b_r1 ky) v
//   flux_t += (f_t0 - i (f_t1 kx + f_t2 ky)) v
// (acc 0/1: the psi or jx / jy right-hand side; flux targets 0..2: psi flux
// x, y or Pi xx, xy, yy).  Derived from the reference's terms by build_desc.
struct ContribCoef {
    double a[2][4];
    double b[2][2];
    double f[3][3];
    // the same coefficients rounded to FP32 (the FP32 form reads these: no
    // conversion per mode)
    float af[2][4];
    float bf[2][2];
    float ff[3][3];
    // nonzero entries: a[r][t] bit 4r + t, b[r][t] bit 8 + 2r + t, f[t][u] bit 12 + 3t + u
    unsigned nz;
};

// coefficient r, t of a contribution in the arithmetic type T
template <typename T>
__host__ __device__ __forceinline__ T coef_a(const ContribCoef& c, int r, int t) {
    if constexpr (sizeof(T) == 4) return c.af[r][t];
    else return c.a[r][t];
}
template <typename T>
__host__ __device__ __forceinline__ T coef_b(const ContribCoef& c, int r, int t) {
    if constexpr (sizeof(T) == 4) return c.bf[r][t];
    else return c.b[r][t];
}
template <typename T>
__host__ __device__ __forceinline__ T coef_f(const ContribCoef& c, int t, int u) {
    if constexpr (sizeof(T) == 4) return c.ff[t][u];
    else return c.f[t][u];
}

struct Group {
    int kind;        // GROUP_PSI or GROUP_MOM
    int alpha;       // order parameter of a PSI group
    int ncontrib;    // forward-transformed products feeding this group
    int contrib[kMaxContrib];   // indices into the product list (U buffer)
    ContribCoef coef[kMaxContrib];
};

struct Source {      // constant spectral source (sine force)
    int field;       // state field index
    int gx, cidx;    // compact mode
    double re, im;
};

struct Desc {
    // geometry: active grid Nax x Nay, compute grid Mx x My; the compact
    // spectral column of gx holds Nay slots c with ky = c (c < Nay/2) or
    // ky = c - Nay (c >= Nay/2); slot Nay/2 (ky = -Nay/2) is always zero.
    int nax, nay, mx, my, hxa;
    int lnay, lhxa;             // log2 of nay, hxa when powers of two, else -1
    int ulg;                    // U planes in blocks of 2^ulg kx columns: element (kx, y)
                                // at ((kx >> ulg) My + y) 2^ulg + (kx mod 2^ulg); 2^ulg divides Nax/2
    int tlg;                    // T planes: -1 row-major (element (kx, y) at y Nax/2 + kx), else
                                // blocks of 2^tlg kx columns laid out as U's; 2^tlg divides Nax/2
    double lx, ly;
    double dkx, dky;            // 2 pi / Lx, 2 pi / Ly
    int N, nfields, mode;

    // inverse column pass: state field of each inverse-transformed slot
    int n_in;
    int in_field[kMaxIn];
    int in_role[kMaxIn];
    double in_scale;            // 1/(Mx*My) of the backward transform

    // row pass
    int make_velocity;          // v = j / rho (compressible) or j / rho0
    int velocity_compressible;
    int n_out;
    OutItem out[kMaxOut];
    double phi4_u[kMaxOp];
    int n_grad;
    int grad_component[8], grad_direction[8];
    double grad_amplitude[8];

    // forward column pass + stage update
    int ngroups;
    Group groups[kMaxGroups];
    double dt, sqrt_dt;
    double mobility[kMaxOp];
    int psi_det_on[kMaxOp];     // psi_det does not return early (src/fcalculator.cc:95-97)
    int psi_lin_on[kMaxOp];
    double mu_k0[kMaxOp], mu_k2[kMaxOp], mu_k4[kMaxOp];
    int mom_det_on;             // j_det does not return early (src/fcalculator.cc:131-133)
    double pressure_coefficient;
    int incompressible_viscous; // eta != 0 in incompressible mode
    double eta, zeta;
    double rho0;
    int n_igrad;                // incompressible force/gradient terms (spectral)
    int igrad_component[8], igrad_direction[8];
    double igrad_amplitude[8];
    int nsources;
    Source sources[kMaxSources];

    // noise (in-kernel Philox; src/fcalculator.cc:496-700)
    int psi_noise_on[kMaxOp];
    double psi_noise_pref[kMaxOp];   // 0.5 * P * M_a, sigma = sqrt(pref * k^2)
    int j_noise_on;
    double j_sigma, j_sqrt_eta, j_sqrt_zeta;
    uint32_t seed_lo, seed_hi;

    FftPlanDev plan_x, plan_y;
    int fixed_y;                // column passes: fixed-length transform where one is built (else plan_y)
};

// Field pointers for one stage (static fields share storage across buffers).
template <typename T>
struct StagePtrs {
    const cx<T>* cur[kMaxFields];
    const cx<T>* base[kMaxFields];
    cx<T>* next[kMaxFields];
    // FP64 copies of the k = 0 mode (FP32 form only; null otherwise)
    const cx<double>* cur_k0;
    const cx<double>* base_k0;
    cx<double>* next_k0;
};

struct StageParams {
    double beta, wb, wc;
    int use_b;              // SRK3: A + beta B; Euler: A only
    uint32_t step_lo, step_hi;   // filled in-kernel from the instance's step counter
    // flux recording (src/buffer_flux.cc:170-194, time_integrator_*.h): weight of
    // this stage's deterministic flux; the first stage clears the step's flux
    // and adds draw A of the noise flux scaled by 1/sqrt(dt); the last stage
    // adds the incompressible projection pressure
    double flux_w;
    int flux_first, flux_last;
};

// Flux request bits (FluxRequest, src/buffer_flux.h)
enum FluxBits : int { FLUX_DENSITY = 1, FLUX_ORDER_PARAMETER = 2, FLUX_MOMENTUM = 4 };

// Everything one instance needs on the device, kept in a device-side table so
// that one launch can serve many instances of different sizes (batched
// ensembles): the workgroups of a launch are mapped to (instance, local block)
// through BlockMap.
template <typename T>
struct InstDev {
    Desc d;
    const cx<T>* twx;         // twiddles exp(-2 pi i j / Mx), j < Mx, then the
                              // radix-8 rows of the fused transforms (fz_tw8_at)
    const cx<T>* twy;         // same for My
    cx<T>* tbuf;              // K1 -> K2 planes, [field] of My rows x Nax/2 kx (Desc::tlg)
    cx<T>* ubuf;              // K2 -> K3 planes, kx-blocked column-major (Desc::ulg)
    int* flag;                // zero-density flag
    const uint64_t* step;     // Philox step counter of this instance
    cx<T>* flux;              // step flux, compact slots [rho x, rho y, (psi x, psi y)*N,
                              // Pi xx, Pi xy, Pi yy]; null unless requested
    int flux_req;             // FluxBits
    int C;                    // kx columns per inverse-column workgroup
    int R;                    // rows per row-pass workgroup
    int ntiles;               // inverse-column tiles = ceil(Nax/2 / C)
    int C3;                   // kx columns per forward-column workgroup
    int ntiles3;              // forward-column tiles
    int k3_order;             // forward-column workgroup order: 1 momentum groups first in each XCD's
                              // range (a launch of few rounds), 0 momentum group first in each tile
};

// An instance whose density reached zero stops at the last step it completed
// (the reference throws inside the failing right-hand side,
// src/fcalculator_workspace.cc:126-128): every later kernel of that instance
// returns at entry, so its state buffers keep the failing step's start.
__device__ __forceinline__ bool instance_failed(const int* flag) {
    return flag != nullptr && *reinterpret_cast<const volatile int*>(flag) != 0;
}

struct BlockMap {
    const int* inst;          // instance of each workgroup of the launch
    const int* first;         // first workgroup of each instance
};

}  // namespace shhd
