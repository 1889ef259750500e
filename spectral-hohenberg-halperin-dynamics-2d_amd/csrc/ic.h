// ic.h — initial conditions evaluated on the device, straight into the
// compact state layout.  Every mode's value is a pure function of
// (seed, channel, gx, ky, draw) through splitmix64 (the reference's mode hash,
// src/initial_condition_equilibrium_gaussian_density.cc:16-50 and
// src/initial_condition_random_vorticity_momentum.cc:16-34), so one lane
// computes one mode with no shared generator state.
#pragma once

#include "../../include/shhd_gpu.h"
#include "desc.h"

namespace shhd {

struct IcDev {
    shhd_initial_condition ic;
    int nay, hxa, mx, my;
    double lx, ly;
    double grid, cell;    // Mx*My and Lx*Ly/(Mx*My)
    double rho0;          // reference density of the state (momentum styles)
    double scale;         // normalization of the random-vorticity / spectrum styles
};

SHHD_HD uint64_t hash_splitmix64(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

SHHD_HD double hash_unit(uint64_t key) {
    return static_cast<double>(hash_splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

// key of the equilibrium styles: seed, channel << 48, gx << 32, ky << 8, draw
SHHD_HD uint64_t mode_key(int seed, int channel, int gx, int ky, int draw) {
    return static_cast<uint64_t>(static_cast<uint32_t>(seed)) ^
           (static_cast<uint64_t>(static_cast<uint32_t>(channel)) << 48) ^
           (static_cast<uint64_t>(static_cast<uint32_t>(gx)) << 32) ^
           (static_cast<uint64_t>(static_cast<uint32_t>(ky)) << 8) ^
           static_cast<uint64_t>(static_cast<uint32_t>(draw));
}

__device__ __forceinline__ cx<double> hashed_gaussian(int seed, int channel, int gx, int ky) {
    const int pky = gx == 0 ? (ky < 0 ? -ky : ky) : ky;
    const double u1 = fmax(hash_unit(mode_key(seed, channel, gx, pky, 0)), 1.0e-300);
    const double u2 = hash_unit(mode_key(seed, channel, gx, pky, 1));
    const double r = sqrt(-2.0 * log(u1));
    const double th = 2.0 * 3.14159265358979323846 * u2;
    const cx<double> v = mk(r * cos(th), r * sin(th));
    return (gx == 0 && ky < 0) ? conj(v) : v;
}

// random phase of random_vorticity (key: seed ^ gx << 32 ^ ky), conjugated on
// the kx = 0, ky < 0 half line
__device__ __forceinline__ cx<double> hashed_phase(int seed, int gx, int ky) {
    const int pky = gx == 0 ? (ky < 0 ? -ky : ky) : ky;
    const uint64_t key = static_cast<uint64_t>(static_cast<uint32_t>(seed)) ^
                         (static_cast<uint64_t>(static_cast<uint32_t>(gx)) << 32) ^
                         static_cast<uint64_t>(static_cast<uint32_t>(pky));
    const double th = 2.0 * 3.14159265358979323846 * hash_unit(key);
    const cx<double> v = mk(cos(th), sin(th));
    return (gx == 0 && ky < 0) ? conj(v) : v;
}

// velocity of the normalized styles before scaling (k != 0)
__device__ __forceinline__ void styled_velocity(const IcDev& d, int gx, int ky, double kx, double kyv, double k2,
                                                cx<double>& ux, cx<double>& uy) {
    const double n = sqrt(static_cast<double>(gx) * gx + static_cast<double>(ky) * ky);
    const cx<double> ph = hashed_phase(d.ic.seed, gx, ky);
    if (d.ic.kind == SHHD_IC_RANDOM_VORTICITY) {
        // src/initial_condition_random_vorticity_momentum.cc:56-99:
        // omega = exp(-((n - n0)/sigma)^2 / 2) e^{i theta}; u = (i ky, -i kx) omega / k^2
        const double z = (n - d.ic.n0) / d.ic.sigma;
        const cx<double> om = exp(-0.5 * z * z) * ph;
        ux = mul_i((kyv / k2) * om);
        uy = mul_mi((kx / k2) * om);
    } else {
        // E(n) ∝ n^4 exp(-2 (n/n0)^2): amplitude n^(3/2) exp(-(n/n0)^2) along
        // i (-ky, kx)/|k| (the factor i keeps the kx = 0 line Hermitian)
        const double a = pow(n, 1.5) * exp(-(n / d.ic.n0) * (n / d.ic.n0));
        const double k = sqrt(k2);
        ux = mul_i((-kyv / k * a) * ph);
        uy = mul_i((kx / k * a) * ph);
    }
}

// value of compact mode (gx, ky) of the target field (ky = -Nay/2 is never active)
__device__ __forceinline__ cx<double> ic_value(const IcDev& d, int gx, int ky) {
    const double two_pi = 2.0 * 3.14159265358979323846;
    const double kx = two_pi * gx / d.lx;
    const double kyv = two_pi * ky / d.ly;
    const double k2 = kx * kx + kyv * kyv;
    const shhd_initial_condition& ic = d.ic;
    if (k2 == 0.0) {
        if (ic.kind == SHHD_IC_EQUILIBRIUM_DENSITY || ic.kind == SHHD_IC_EQUILIBRIUM_ORDER_PARAMETER)
            return mk(ic.value * d.grid, 0.0);
        return czero<double>();
    }
    switch (ic.kind) {
        case SHHD_IC_EQUILIBRIUM_DENSITY: {
            if (ic.kBT == 0.0) return czero<double>();
            const double sigma = sqrt(ic.kBT * d.grid * ic.value / (2.0 * ic.pressure_coefficient * d.cell));
            return sigma * hashed_gaussian(ic.seed, 0, gx, ky);
        }
        case SHHD_IC_EQUILIBRIUM_ORDER_PARAMETER: {
            if (ic.kBT == 0.0) return czero<double>();
            const double coef = ic.k0 + ic.k2 * k2 + ic.k4 * k2 * k2;   // > 0, host-checked
            const double sigma = sqrt(ic.kBT * d.grid / (2.0 * coef * d.cell));
            return sigma * hashed_gaussian(ic.seed, ic.component, gx, ky);
        }
        case SHHD_IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE: {
            if (ic.kBT == 0.0) return czero<double>();
            const double sigma = sqrt(ic.kBT * d.grid * d.rho0 / (2.0 * d.cell));
            return sigma * hashed_gaussian(ic.seed, ic.component, gx, ky);
        }
        case SHHD_IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE: {
            if (ic.kBT == 0.0) return czero<double>();
            const double sigma = sqrt(ic.kBT * d.grid * d.rho0 / (2.0 * d.cell));
            const double k = sqrt(k2);
            const double c = ic.component == 0 ? -kyv / k : kx / k;
            return c * (sigma * hashed_gaussian(ic.seed, 0, gx, ky));
        }
        default: {
            cx<double> ux, uy;
            styled_velocity(d, gx, ky, kx, kyv, k2, ux, uy);
            return (d.rho0 * d.scale) * (ic.component == 0 ? ux : uy);
        }
    }
}

// Write every compact slot of the target field.
template <typename T>
__global__ __launch_bounds__(256) void k_initial_condition(IcDev d, cx<T>* __restrict__ field) {
    const size_t total = static_cast<size_t>(d.hxa) * d.nay;
    const int half = d.nay / 2;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int gx = static_cast<int>(i / d.nay);
        const int c = static_cast<int>(i - static_cast<size_t>(gx) * d.nay);
        cx<double> v = czero<double>();
        if (c != half) v = ic_value(d, gx, c < half ? c : c - d.nay);
        field[i] = mk(static_cast<T>(v.re), static_cast<T>(v.im));
    }
}

// Per-workgroup partial sums of weight * (|ux|^2 + |uy|^2) (weight 1 on the
// kx = 0 line, 2 elsewhere) for the normalization of the velocity styles;
// the host adds the partials in order, so the sum is reproducible.
__global__ __launch_bounds__(256) void k_ic_norm(IcDev d, double* __restrict__ partials) {
    __shared__ double red[256];
    const size_t total = static_cast<size_t>(d.hxa) * d.nay;
    const int half = d.nay / 2;
    const double two_pi = 2.0 * 3.14159265358979323846;
    double acc = 0.0;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int gx = static_cast<int>(i / d.nay);
        const int c = static_cast<int>(i - static_cast<size_t>(gx) * d.nay);
        if (c == half) continue;
        const int ky = c < half ? c : c - d.nay;
        const double kx = two_pi * gx / d.lx;
        const double kyv = two_pi * ky / d.ly;
        const double k2 = kx * kx + kyv * kyv;
        if (k2 == 0.0) continue;
        cx<double> ux, uy;
        styled_velocity(d, gx, ky, kx, kyv, k2, ux, uy);
        acc += (gx == 0 ? 1.0 : 2.0) * (ux.re * ux.re + ux.im * ux.im + uy.re * uy.re + uy.im * uy.im);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (static_cast<int>(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

}  // namespace shhd
