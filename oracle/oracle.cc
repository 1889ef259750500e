// oracle.cc — FP64 CPU restatement of the reference solver's RHS and time-step path.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Each routine cites the reference
// file:line it restates.  The FFT is this file's own mixed-radix Stockham
// transform (the reference delegates to heFFTe 2.4.x / FFTW 3.3.10, neither of
// which is vendored); the tests pin it against the DFT definition.
//
// Layout and conventions (src/state.cc:8-38, src/domain.cc:139-149,
// src/fourier_transform.h:43,50): spectra are complex[(Mx/2+1)*My] per field,
// index gy*(Mx/2+1)+gx; physical arrays are double[My*Mx], index gy*Mx+gx;
// forward is unscaled with exp(-i k x), backward carries 1/(Mx*My).

#include "oracle.h"

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cplx = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;
const cplx I(0.0, 1.0);

// -------------------------------------------------------------------------
// 1D complex FFT: Stockham autosort, radices 4, 2, 3, 5 and a generic DFT for
// any other prime factor.  exp(sign * 2 pi i j k / n), unscaled.
// -------------------------------------------------------------------------
struct FftPlan {
    int n = 0;
    std::vector<int> radices;
    std::vector<cplx> w;   // w[j] = exp(-2 pi i j / n)

    explicit FftPlan(int n_) : n(n_) {
        if (n <= 0) throw std::runtime_error("FFT length must be positive");
        int m = n;
        while (m % 4 == 0) { radices.push_back(4); m /= 4; }
        while (m % 2 == 0) { radices.push_back(2); m /= 2; }
        while (m % 3 == 0) { radices.push_back(3); m /= 3; }
        while (m % 5 == 0) { radices.push_back(5); m /= 5; }
        for (int p = 7; m > 1; p += 2) {
            while (m % p == 0) { radices.push_back(p); m /= p; }
        }
        w.resize(static_cast<size_t>(n));
        for (int j = 0; j < n; ++j) {
            // reduce to the first octant for accuracy
            const long long jj = j;
            const double a = 2.0 * kPi * static_cast<double>(jj) / static_cast<double>(n);
            w[static_cast<size_t>(j)] = cplx(std::cos(a), -std::sin(a));
        }
    }

    cplx tw(long long e, int sign) const {
        long long r = e % n;
        if (r < 0) r += n;
        const cplx v = w[static_cast<size_t>(r)];
        return sign < 0 ? v : std::conj(v);
    }

    // x is input (destroyed), y is scratch; result returned in the pointer
    // that holds it after the passes.
    cplx* run(cplx* x, cplx* y, int sign) const {
        int ns = 1;
        cplx v[64];
        cplx o[64];
        for (int r : radices) {
            const int stride = n / r;
            for (int j = 0; j < stride; ++j) {
                const int k = j % ns;
                const long long tstep = static_cast<long long>(n / (ns * r));
                for (int q = 0; q < r; ++q) {
                    const cplx a = x[j + q * stride];
                    v[q] = (q == 0 || k == 0) ? a : a * tw(static_cast<long long>(q) * k * tstep, sign);
                }
                if (r == 2) {
                    o[0] = v[0] + v[1];
                    o[1] = v[0] - v[1];
                } else if (r == 4) {
                    const cplx a0 = v[0] + v[2], a1 = v[0] - v[2];
                    const cplx b0 = v[1] + v[3], b1 = v[1] - v[3];
                    const cplx rot = (sign < 0) ? cplx(b1.imag(), -b1.real()) : cplx(-b1.imag(), b1.real());
                    o[0] = a0 + b0;
                    o[2] = a0 - b0;
                    o[1] = a1 + rot;
                    o[3] = a1 - rot;
                } else {
                    for (int p = 0; p < r; ++p) {
                        cplx acc = v[0];
                        for (int q = 1; q < r; ++q) {
                            acc += v[q] * tw(static_cast<long long>(p) * q * (n / r), sign);
                        }
                        o[p] = acc;
                    }
                }
                const int base = (j / ns) * ns * r + k;
                for (int p = 0; p < r; ++p) y[base + p * ns] = o[p];
            }
            std::swap(x, y);
            ns *= r;
        }
        return x;
    }
};

std::mutex g_plan_mutex;
std::map<int, std::shared_ptr<FftPlan>> g_plans;

std::shared_ptr<FftPlan> plan_for(int n) {
    std::lock_guard<std::mutex> lock(g_plan_mutex);
    auto it = g_plans.find(n);
    if (it != g_plans.end()) return it->second;
    auto p = std::make_shared<FftPlan>(n);
    g_plans[n] = p;
    return p;
}

void fft_inplace(const FftPlan& plan, cplx* data, std::vector<cplx>& scratch, int sign) {
    scratch.resize(static_cast<size_t>(plan.n));
    cplx* res = plan.run(data, scratch.data(), sign);
    if (res != data) std::memcpy(data, res, sizeof(cplx) * static_cast<size_t>(plan.n));
}

// -------------------------------------------------------------------------
// 2D real transforms with the reference's normalization
// (src/fourier_transform.h:39-51: forward scale::none, backward scale::full).
// Forward: R2C along x (direction 0, src/domain.h:17), then C2C along y.
// Backward: C2C along y, then C2R along x; the C2R reads the half spectrum
// as Hermitian (imaginary parts of the kx=0 and Nyquist entries ignored).
// -------------------------------------------------------------------------
void fft2_r2c_many(int mx, int my, int nfields, const double* phys, cplx* spec) {
    const int hx = mx / 2 + 1;
    auto px = plan_for(mx);
    auto py = plan_for(my);
    for (int f = 0; f < nfields; ++f) {
        const double* in = phys + static_cast<size_t>(f) * mx * my;
        cplx* out = spec + static_cast<size_t>(f) * hx * my;
#pragma omp parallel
        {
            std::vector<cplx> row(static_cast<size_t>(mx)), scratch;
#pragma omp for schedule(static)
            for (int y = 0; y < my; ++y) {
                for (int x = 0; x < mx; ++x) row[x] = cplx(in[static_cast<size_t>(y) * mx + x], 0.0);
                fft_inplace(*px, row.data(), scratch, -1);
                for (int x = 0; x < hx; ++x) out[static_cast<size_t>(y) * hx + x] = row[x];
            }
        }
#pragma omp parallel
        {
            std::vector<cplx> col(static_cast<size_t>(my)), scratch;
#pragma omp for schedule(static)
            for (int x = 0; x < hx; ++x) {
                for (int y = 0; y < my; ++y) col[y] = out[static_cast<size_t>(y) * hx + x];
                fft_inplace(*py, col.data(), scratch, -1);
                for (int y = 0; y < my; ++y) out[static_cast<size_t>(y) * hx + x] = col[y];
            }
        }
    }
}

void fft2_c2r_many(int mx, int my, int nfields, const cplx* spec, double* phys, std::vector<cplx>& work) {
    const int hx = mx / 2 + 1;
    auto px = plan_for(mx);
    auto py = plan_for(my);
    const double scale = 1.0 / (static_cast<double>(mx) * static_cast<double>(my));
    work.resize(static_cast<size_t>(hx) * my);
    for (int f = 0; f < nfields; ++f) {
        const cplx* in = spec + static_cast<size_t>(f) * hx * my;
        double* out = phys + static_cast<size_t>(f) * mx * my;
        cplx* tmp = work.data();
#pragma omp parallel
        {
            std::vector<cplx> col(static_cast<size_t>(my)), scratch;
#pragma omp for schedule(static)
            for (int x = 0; x < hx; ++x) {
                for (int y = 0; y < my; ++y) col[y] = in[static_cast<size_t>(y) * hx + x];
                fft_inplace(*py, col.data(), scratch, +1);
                for (int y = 0; y < my; ++y) tmp[static_cast<size_t>(y) * hx + x] = col[y];
            }
        }
#pragma omp parallel
        {
            std::vector<cplx> row(static_cast<size_t>(mx)), scratch;
#pragma omp for schedule(static)
            for (int y = 0; y < my; ++y) {
                const cplx* s = tmp + static_cast<size_t>(y) * hx;
                row[0] = cplx(s[0].real(), 0.0);
                for (int x = 1; x < hx; ++x) row[x] = s[x];
                if (mx % 2 == 0) row[mx / 2] = cplx(s[mx / 2].real(), 0.0);
                for (int x = hx; x < mx; ++x) row[x] = std::conj(s[mx - x]);
                fft_inplace(*px, row.data(), scratch, +1);
                for (int x = 0; x < mx; ++x) out[static_cast<size_t>(y) * mx + x] = row[x].real() * scale;
            }
        }
    }
}

// -------------------------------------------------------------------------
// Spectral mask (src/spectral_mask.cc:7-72) on a single-rank domain.
// -------------------------------------------------------------------------
// -------------------------------------------------------------------------
// Mode-hash random numbers of the initial conditions
// (src/initial_condition_equilibrium_gaussian_density.cc:12-50,
//  src/initial_condition_random_vorticity_momentum.cc:12-34)
// -------------------------------------------------------------------------
uint64_t ic_splitmix64(uint64_t v) {
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

double ic_uniform(int seed, int channel, int gx, int ky, int draw) {
    uint64_t key = static_cast<uint64_t>(static_cast<uint32_t>(seed));
    key ^= static_cast<uint64_t>(static_cast<uint32_t>(channel)) << 48;
    key ^= static_cast<uint64_t>(static_cast<uint32_t>(gx)) << 32;
    key ^= static_cast<uint64_t>(static_cast<uint32_t>(ky)) << 8;
    key ^= static_cast<uint64_t>(static_cast<uint32_t>(draw));
    return static_cast<double>(ic_splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

// gaussian_mode: Box-Muller of two hashed uniforms; the kx = 0 line hashes
// |ky| and conjugates ky < 0
cplx ic_gaussian(int seed, int channel, int gx, int ky) {
    const int phase_ky = gx == 0 ? std::abs(ky) : ky;
    const double u1 = std::max(ic_uniform(seed, channel, gx, phase_ky, 0), 1.0e-300);
    const double u2 = ic_uniform(seed, channel, gx, phase_ky, 1);
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double th = 2.0 * kPi * u2;
    cplx v(r * std::cos(th), r * std::sin(th));
    return (gx == 0 && ky < 0) ? std::conj(v) : v;
}

// random_phase of random_vorticity: one hashed uniform without channel/draw
cplx ic_phase(int seed, int gx, int ky) {
    uint64_t key = static_cast<uint64_t>(static_cast<uint32_t>(seed));
    key ^= static_cast<uint64_t>(static_cast<uint32_t>(gx)) << 32;
    key ^= static_cast<uint64_t>(static_cast<uint32_t>(ky));
    const double u = static_cast<double>(ic_splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
    const double th = 2.0 * kPi * u;
    return cplx(std::cos(th), std::sin(th));
}

struct Mode {
    size_t index;
    int gx, gy;
    double kx, ky, k2;
};

enum class Dyn { Compressible, Incompressible, Quiescent };

struct Oracle {
    shhd_config cfg;
    std::string error;

    int N = 0;
    int mx = 0, my = 0, hx = 0;
    size_t spec = 0, phys = 0;
    int nfields = 0;
    Dyn dyn = Dyn::Compressible;
    bool srk3 = true;
    double dt = 0.0, sqrt_dt = 0.0;

    std::vector<Mode> active;
    std::vector<size_t> inactive;

    // PhysicalRHSFieldRequests / PhysicalFieldLayout
    // (src/fcalculator_field_requests_and_layout.h:30-221)
    bool transform_psi = false, transform_rho = false, transform_j = false, make_velocity = false;
    int off_psi = -1, off_rho = -1, off_jx = -1, off_jy = -1, off_vx = -1, off_vy = -1;
    int ntransform = 0, nphysical = 0;

    bool fe_has_physical_mu = false;
    bool has_psi_sto = false, has_j_sto = false;

    std::vector<cplx> state;          // u
    std::vector<cplx> u_old, u1, u2, det, sto_a, sto_b, jinc_x, jinc_y;

    // workspace (src/fcalculator_workspace.cc:35-38, 195-227)
    std::vector<cplx> spectral_pack;
    std::vector<double> physical_pack;
    std::vector<double> tmp_phys;
    std::vector<cplx> tmp_spec;
    std::vector<cplx> c2r_work;
    const cplx* cached_state = nullptr;
    double cached_time = 0.0;
    bool physical_ready = false;
    bool rho0_ready = false;
    double rho0 = 0.0;

    std::mt19937_64 rng;

    // FluxBuffer (src/buffer_flux.{h,cc}): slots [rho x, rho y, (psi x, psi y) per
    // order parameter, Pi xx, Pi xy, Pi yy], complex[spec] each
    bool req_density = false, req_op = false, req_mom = false;
    std::vector<cplx> flux_total, flux_det, flux_sto;
    std::vector<cplx>* rec = nullptr;   // buffer the RHS terms record into
    int nslots() const { return 2 + 2 * N + 3; }
    bool flux_requested() const { return req_density || req_op || req_mom; }
    cplx* fslot(int k) { return rec->data() + static_cast<size_t>(k) * spec; }
    cplx* f_rho(int c) { return fslot(c); }
    cplx* f_op(int a, int c) { return fslot(2 + 2 * a + c); }
    cplx* f_pi(int c) { return fslot(2 + 2 * N + c); }   // 0 xx, 1 xy, 2 yy
    std::normal_distribution<double> normal{0.0, 1.0};

    double kx_of(int gx) const { return 2.0 * kPi * static_cast<double>(gx) / cfg.lx; }
    double ky_of(int gy) const {
        const int w = (gy <= my / 2) ? gy : gy - my;
        return 2.0 * kPi * static_cast<double>(w) / cfg.ly;
    }
    // SpectralMask2D::is_active_mode (src/spectral_mask.cc:45-52)
    bool is_active(int gx, int gy) const {
        if (gx < 0 || gx >= cfg.nx_active / 2) return false;
        const int ky = (gy <= my / 2) ? gy : gy - my;
        return std::abs(ky) < cfg.ny_active / 2;
    }

    cplx* field(std::vector<cplx>& s, int f) { return s.data() + static_cast<size_t>(f) * spec; }
    const cplx* field(const cplx* s, int f) const { return s + static_cast<size_t>(f) * spec; }
    int psi_field(int a) const { return 1 + a; }
    int jx_field() const { return N + 1; }
    int jy_field() const { return N + 2; }

    explicit Oracle(const shhd_config& c) : cfg(c) {
        N = cfg.num_order_parameters;
        if (N < 0 || N > SHHD_MAX_ORDER_PARAMETERS) throw std::runtime_error("num_order_parameters out of range");
        if (cfg.nx_active <= 0 || cfg.ny_active <= 0 || cfg.nx_active % 2 || cfg.ny_active % 2)
            throw std::runtime_error("grid must be even.");
        if (cfg.nx < cfg.nx_active || cfg.ny < cfg.ny_active)
            throw std::runtime_error("SpectralMask2D active grid cannot exceed compute grid.");
        if (!(cfg.lx > 0.0) || !(cfg.ly > 0.0)) throw std::runtime_error("domain lengths must be positive");
        if (!(cfg.dt > 0.0)) throw std::runtime_error("timestep must be positive");
        mx = cfg.nx; my = cfg.ny; hx = mx / 2 + 1;
        spec = static_cast<size_t>(hx) * my;
        phys = static_cast<size_t>(mx) * my;
        nfields = N + 3;
        if (cfg.mode == SHHD_MODE_COMPRESSIBLE) dyn = Dyn::Compressible;
        else if (cfg.mode == SHHD_MODE_INCOMPRESSIBLE) dyn = Dyn::Incompressible;
        else if (cfg.mode == SHHD_MODE_QUIESCENT) dyn = Dyn::Quiescent;
        else throw std::runtime_error("unknown time_evolution type");
        if (cfg.integrator != SHHD_INTEGRATOR_EULER && cfg.integrator != SHHD_INTEGRATOR_SRK3)
            throw std::runtime_error("unknown time_evolution type");
        srk3 = cfg.integrator == SHHD_INTEGRATOR_SRK3;
        dt = cfg.dt;
        sqrt_dt = std::sqrt(dt);

        if (cfg.eta < 0.0 || cfg.zeta < 0.0) throw std::runtime_error("transport coefficient requires nonnegative eta.");
        for (int a = 0; a < N; ++a) {
            if (cfg.mobility[a] < 0.0) throw std::runtime_error("transport coefficient requires nonnegative order parameter mobility.");
        }
        if (cfg.free_energy_kind == SHHD_FREE_ENERGY_PHI4) {
            for (int a = 0; a < N; ++a) {
                if (cfg.phi4_u[a] < 0.0) throw std::runtime_error("phi4 free energy requires nonnegative u coefficient.");
                if (cfg.phi4_u[a] != 0.0) fe_has_physical_mu = true;
            }
        }
        // ParamParser::validate_configuration (src/param_parser.cc:771-805)
        const bool quiescent = dyn == Dyn::Quiescent;
        if (cfg.order_parameter_advection && N == 0)
            throw std::runtime_error("fix nonlinear order_parameter requires at least one order parameter.");
        if (quiescent && cfg.momentum_advection)
            throw std::runtime_error("quiescent time evolution cannot use momentum nonlinear advection.");
        if (quiescent && cfg.order_parameter_advection)
            throw std::runtime_error("quiescent time evolution cannot use order_parameter nonlinear advection.");
        for (int i = 0; i < cfg.num_sine_forces; ++i)
            if (quiescent && cfg.sine_forces[i].momentum)
                throw std::runtime_error("quiescent time evolution cannot use momentum force/sine.");
        if (quiescent && cfg.num_gradient_forces > 0)
            throw std::runtime_error("quiescent time evolution cannot use force/gradient.");

        // mask (src/spectral_mask.cc:7-38)
        size_t index = 0;
        for (int gy = 0; gy < my; ++gy) {
            const double ky = ky_of(gy);
            for (int gx = 0; gx < hx; ++gx) {
                const double kx = kx_of(gx);
                if (is_active(gx, gy)) active.push_back({index, gx, gy, kx, ky, kx * kx + ky * ky});
                else inactive.push_back(index);
                ++index;
            }
        }

        // requests (src/fcalculator_field_requests_and_layout.h:34-127)
        bool need_psi = false, need_rho = false, need_j = false, need_vel = false;
        if (N > 0) {
            if (fe_has_physical_mu) need_psi = true;
            if (cfg.order_parameter_advection) { need_psi = true; need_j = true; need_vel = true; }
            if (cfg.num_gradient_forces > 0 && dyn == Dyn::Compressible) { need_j = true; need_vel = true; }
        }
        if (!quiescent) {
            const bool has_visc = cfg.eta != 0.0 || cfg.zeta != 0.0;
            if (dyn == Dyn::Compressible && has_visc) { need_j = true; need_vel = true; }
            if (cfg.momentum_advection) { need_j = true; need_vel = true; }
        }
        if (quiescent && (need_rho || need_j || need_vel))
            throw std::runtime_error("quiescent workspace cannot request hydrodynamic physical fields.");
        transform_psi = need_psi;
        transform_j = need_j;
        make_velocity = need_vel;
        transform_rho = need_rho || (dyn == Dyn::Compressible && need_vel);
        // PhysicalFieldLayout::assign_offsets (…layout.h:196-220)
        if (transform_psi) { off_psi = ntransform; ntransform += N; }
        if (transform_rho) { off_rho = ntransform; ntransform += 1; }
        if (transform_j) { off_jx = ntransform; off_jy = ntransform + 1; ntransform += 2; }
        nphysical = ntransform;
        if (make_velocity) { off_vx = nphysical; off_vy = nphysical + 1; nphysical += 2; }

        // noise enablement (src/fluid_solver.h:79-94, src/param_parser.cc:628-638)
        has_psi_sto = cfg.noise_order_parameter && N > 0;
        has_j_sto = cfg.noise_momentum && !quiescent;

        const size_t total = static_cast<size_t>(nfields) * spec;
        state.assign(total, cplx(0, 0));
        u_old = u1 = u2 = det = sto_a = sto_b = state;
        jinc_x.assign(spec, cplx(0, 0));
        jinc_y.assign(spec, cplx(0, 0));
        spectral_pack.assign(static_cast<size_t>(ntransform) * spec, cplx(0, 0));
        physical_pack.assign(static_cast<size_t>(nphysical) * phys, 0.0);
        flux_total.assign(static_cast<size_t>(nslots()) * spec, cplx(0, 0));
        flux_det = flux_sto = flux_total;
        const int cap = std::max(N, 9);   // src/fcalculator.cc:11-19
        tmp_phys.assign(static_cast<size_t>(cap) * phys, 0.0);
        tmp_spec.assign(static_cast<size_t>(cap) * spec, cplx(0, 0));

        // FCalculator ctor seeding (src/fcalculator.cc:49-53), single rank
        std::seed_seq seq{static_cast<std::uint32_t>(static_cast<int>(cfg.seed)), static_cast<std::uint32_t>(0)};
        rng.seed(seq);
    }

    // ---------------------------------------------------------------------
    // FCalculatorWorkspace (src/fcalculator_workspace.cc:41-145)
    // ---------------------------------------------------------------------
    double reference_density(const cplx* s) {
        if (rho0_ready) return rho0;
        const double grid = static_cast<double>(mx) * static_cast<double>(my);
        rho0 = s[0].real() / grid;   // rho_hat(0,0) / (Mx*My)
        if (rho0 == 0.0) throw std::runtime_error("reference density must be nonzero.");
        rho0_ready = true;
        return rho0;
    }

    void ensure_physical_fields(const cplx* s, double time) {
        if (cached_state == s && cached_time == time && physical_ready) return;
        cached_state = s;
        cached_time = time;
        physical_ready = false;
        std::fill(spectral_pack.begin(), spectral_pack.end(), cplx(0, 0));
        std::fill(physical_pack.begin(), physical_pack.end(), 0.0);
        if (transform_psi)
            std::memcpy(spectral_pack.data() + static_cast<size_t>(off_psi) * spec, field(s, 1), sizeof(cplx) * N * spec);
        if (transform_rho)
            std::memcpy(spectral_pack.data() + static_cast<size_t>(off_rho) * spec, field(s, 0), sizeof(cplx) * spec);
        if (transform_j) {
            std::memcpy(spectral_pack.data() + static_cast<size_t>(off_jx) * spec, field(s, jx_field()), sizeof(cplx) * spec);
            std::memcpy(spectral_pack.data() + static_cast<size_t>(off_jy) * spec, field(s, jy_field()), sizeof(cplx) * spec);
        }
        if (ntransform > 0) fft2_c2r_many(mx, my, ntransform, spectral_pack.data(), physical_pack.data(), c2r_work);
        if (make_velocity) {
            const double* jx = physical_pack.data() + static_cast<size_t>(off_jx) * phys;
            const double* jy = physical_pack.data() + static_cast<size_t>(off_jy) * phys;
            double* vx = physical_pack.data() + static_cast<size_t>(off_vx) * phys;
            double* vy = physical_pack.data() + static_cast<size_t>(off_vy) * phys;
            if (dyn == Dyn::Compressible) {
                const double* rho = physical_pack.data() + static_cast<size_t>(off_rho) * phys;
                bool zero = false;
#pragma omp parallel for reduction(|| : zero)
                for (long long i = 0; i < static_cast<long long>(phys); ++i) {
                    if (rho[i] == 0.0) { zero = true; continue; }
                    vx[i] = jx[i] / rho[i];
                    vy[i] = jy[i] / rho[i];
                }
                if (zero) throw std::runtime_error("compressible velocity encountered zero density.");
            } else if (dyn == Dyn::Incompressible) {
                const double r0 = reference_density(s);
#pragma omp parallel for
                for (long long i = 0; i < static_cast<long long>(phys); ++i) {
                    vx[i] = jx[i] / r0;
                    vy[i] = jy[i] / r0;
                }
            }
        }
        physical_ready = true;
    }

    double* temporary_physical(int n) {
        std::fill(tmp_phys.begin(), tmp_phys.begin() + static_cast<size_t>(n) * phys, 0.0);
        return tmp_phys.data();
    }
    cplx* temporary_spectral(int n) {
        std::fill(tmp_spec.begin(), tmp_spec.begin() + static_cast<size_t>(n) * spec, cplx(0, 0));
        return tmp_spec.data();
    }
    const double* phys_field(int off) const { return physical_pack.data() + static_cast<size_t>(off) * phys; }

    // ---------------------------------------------------------------------
    // FCalculator deterministic parts (src/fcalculator.cc:60-494, 702-840)
    // ---------------------------------------------------------------------
    void rho_det(const cplx* s, cplx* out) {          // fcalculator.cc:60-84
        std::fill(out, out + spec, cplx(0, 0));
        const cplx* jx = field(s, jx_field());
        const cplx* jy = field(s, jy_field());
        for (const Mode& m : active) out[m.index] = -I * (m.kx * jx[m.index] + m.ky * jy[m.index]);
        if (rec && req_density) {
            for (const Mode& m : active) {
                f_rho(0)[m.index] += jx[m.index];
                f_rho(1)[m.index] += jy[m.index];
            }
        }
    }

    void psi_det(int a, const cplx* s, cplx* out, double t) {   // fcalculator.cc:86-108
        std::fill(out, out + spec, cplx(0, 0));
        const double mob = cfg.mobility[a];
        const bool adv = cfg.order_parameter_advection != 0;
        if (mob == 0.0 && !adv) return;
        if (mob != 0.0) add_psi_linear(a, mob, s, out);
        if (adv || fe_has_physical_mu) add_psi_physical(a, mob, s, out, t);
        add_psi_sine_force(a, out);
        add_psi_gradient_force(a, s, out, t);
    }

    void add_psi_linear(int a, double mob, const cplx* s, cplx* out) {   // fcalculator.cc:152-187
        const double c0 = cfg.mu_k0[a], c2 = cfg.mu_k2[a], c4 = cfg.mu_k4[a];
        if (c0 == 0.0 && c2 == 0.0 && c4 == 0.0) return;
        const cplx* psi = field(s, psi_field(a));
        for (const Mode& m : active) {
            const double k2 = m.k2;
            const cplx mu = (c0 + c2 * k2 + c4 * k2 * k2) * psi[m.index];
            const cplx fx = -cplx(0.0, mob * m.kx) * mu;
            const cplx fy = -cplx(0.0, mob * m.ky) * mu;
            out[m.index] += -I * (m.kx * fx + m.ky * fy);
            if (rec && req_op) {
                f_op(a, 0)[m.index] += fx;
                f_op(a, 1)[m.index] += fy;
            }
        }
    }

    void add_psi_physical(int a, double mob, const cplx* s, cplx* out, double t) {   // fcalculator.cc:189-274
        const bool phys_mu = mob != 0.0 && fe_has_physical_mu;
        const bool adv = cfg.order_parameter_advection != 0;
        ensure_physical_fields(s, t);
        int nf = 0;
        const int f_mu = phys_mu ? nf++ : -1;
        const int f_x = adv ? nf++ : -1;
        const int f_y = adv ? nf++ : -1;
        double* pf = temporary_physical(nf);
        if (phys_mu) {
            const double* psi = phys_field(off_psi) + static_cast<size_t>(a) * phys;
            double* mu = pf + static_cast<size_t>(f_mu) * phys;
            const double u = cfg.phi4_u[a];
            // Phi4FreeEnergy::physical_chemical_potential (src/model_free_energy_phi4.cc:47-55)
#pragma omp parallel for
            for (long long i = 0; i < static_cast<long long>(phys); ++i) {
                const double v = psi[i];
                mu[i] = u * v * v * v;
            }
        }
        if (adv) {
            const double* psi = phys_field(off_psi) + static_cast<size_t>(a) * phys;
            const double* vx = phys_field(off_vx);
            const double* vy = phys_field(off_vy);
            double* fx = pf + static_cast<size_t>(f_x) * phys;
            double* fy = pf + static_cast<size_t>(f_y) * phys;
#pragma omp parallel for
            for (long long i = 0; i < static_cast<long long>(phys); ++i) {
                fx[i] = psi[i] * vx[i];
                fy[i] = psi[i] * vy[i];
            }
        }
        cplx* sf = temporary_spectral(nf);
        if (nf > 0) fft2_r2c_many(mx, my, nf, pf, sf);
        if (phys_mu) {
            const cplx* mu = sf + static_cast<size_t>(f_mu) * spec;
            for (const Mode& m : active) {
                const cplx fx = -cplx(0.0, mob * m.kx) * mu[m.index];
                const cplx fy = -cplx(0.0, mob * m.ky) * mu[m.index];
                out[m.index] += -I * (m.kx * fx + m.ky * fy);
                if (rec && req_op) {
                    f_op(a, 0)[m.index] += fx;
                    f_op(a, 1)[m.index] += fy;
                }
            }
        }
        if (adv) {
            const cplx* fx = sf + static_cast<size_t>(f_x) * spec;
            const cplx* fy = sf + static_cast<size_t>(f_y) * spec;
            for (const Mode& m : active) {
                out[m.index] += -I * (m.kx * fx[m.index] + m.ky * fy[m.index]);
                if (rec && req_op) {
                    f_op(a, 0)[m.index] += fx[m.index];
                    f_op(a, 1)[m.index] += fy[m.index];
                }
            }
        }
    }

    void j_det(const cplx* s, cplx* ox, cplx* oy, double t) {   // fcalculator.cc:110-150
        std::fill(ox, ox + spec, cplx(0, 0));
        std::fill(oy, oy + spec, cplx(0, 0));
        if (dyn == Dyn::Quiescent) return;
        const double pc = cfg.pressure_coefficient;
        const double eta = cfg.eta, zeta = cfg.zeta;
        const bool lin_p = pc != 0.0;
        const bool visc = eta != 0.0 || zeta != 0.0;
        const bool adv = cfg.momentum_advection != 0;
        if (!lin_p && !visc && !adv) return;
        if (lin_p) add_pressure(pc, s, ox, oy);
        if (visc && dyn == Dyn::Incompressible) add_incompressible_viscous(eta, s, ox, oy);
        if (visc && dyn == Dyn::Compressible) add_compressible_viscous(eta, zeta, s, ox, oy, t);
        if (adv) add_momentum_physical(s, ox, oy, t);
        add_momentum_sine_force(ox, oy);
    }

    void add_pressure(double pc, const cplx* s, cplx* ox, cplx* oy) {   // fcalculator.cc:276-300
        const cplx* rho = field(s, 0);
        for (const Mode& m : active) {
            const cplx p = pc * rho[m.index];
            ox[m.index] += -cplx(0.0, m.kx) * p;
            oy[m.index] += -cplx(0.0, m.ky) * p;
            if (rec && req_mom) {
                f_pi(0)[m.index] += p;
                f_pi(2)[m.index] += p;
            }
        }
    }

    void add_incompressible_viscous(double eta, const cplx* s, cplx* ox, cplx* oy) {   // fcalculator.cc:302-341
        const double r0 = reference_density(s);
        const cplx* jx = field(s, jx_field());
        const cplx* jy = field(s, jy_field());
        for (const Mode& m : active) {
            if (eta != 0.0) {
                const cplx vx = jx[m.index] / r0;
                const cplx vy = jy[m.index] / r0;
                const cplx pxx = -cplx(0.0, eta) * (2.0 * m.kx * vx);
                const cplx pxy = -cplx(0.0, eta) * (m.ky * vx + m.kx * vy);
                const cplx pyy = -cplx(0.0, eta) * (2.0 * m.ky * vy);
                ox[m.index] += -I * (m.kx * pxx + m.ky * pxy);
                oy[m.index] += -I * (m.kx * pxy + m.ky * pyy);
                if (rec && req_mom) {
                    f_pi(0)[m.index] += pxx;
                    f_pi(1)[m.index] += pxy;
                    f_pi(2)[m.index] += pyy;
                }
            }
        }
    }

    void add_compressible_viscous(double eta, double zeta, const cplx* s, cplx* ox, cplx* oy, double t) {   // fcalculator.cc:343-398
        ensure_physical_fields(s, t);
        double* pf = temporary_physical(2);
        std::memcpy(pf, phys_field(off_vx), sizeof(double) * phys);
        std::memcpy(pf + phys, phys_field(off_vy), sizeof(double) * phys);
        cplx* sf = temporary_spectral(2);
        fft2_r2c_many(mx, my, 2, pf, sf);
        const cplx* vx = sf;
        const cplx* vy = sf + spec;
        for (const Mode& m : active) {
            const size_t i = m.index;
            const cplx kv = m.kx * vx[i] + m.ky * vy[i];
            const cplx pxx = -cplx(0.0, 1) * (2.0 * eta * m.kx * vx[i] - (eta - zeta) * kv);
            const cplx pxy = -cplx(0.0, eta) * (m.ky * vx[i] + m.kx * vy[i]);
            const cplx pyy = -cplx(0.0, 1) * (2.0 * eta * m.ky * vy[i] - (eta - zeta) * kv);
            ox[i] += -I * (m.kx * pxx + m.ky * pxy);
            oy[i] += -I * (m.kx * pxy + m.ky * pyy);
            if (rec && req_mom) {
                f_pi(0)[i] += pxx;
                f_pi(1)[i] += pxy;
                f_pi(2)[i] += pyy;
            }
        }
    }

    void add_momentum_physical(const cplx* s, cplx* ox, cplx* oy, double t) {   // fcalculator.cc:400-494
        ensure_physical_fields(s, t);
        double* pf = temporary_physical(3);
        const double* jx = phys_field(off_jx);
        const double* jy = phys_field(off_jy);
        const double* vx = phys_field(off_vx);
        const double* vy = phys_field(off_vy);
        double* pxx = pf;
        double* pxy = pf + phys;
        double* pyy = pf + 2 * phys;
#pragma omp parallel for
        for (long long i = 0; i < static_cast<long long>(phys); ++i) {
            pxx[i] = jx[i] * vx[i];
            pxy[i] = jx[i] * vy[i];
            pyy[i] = jy[i] * vy[i];
        }
        cplx* sf = temporary_spectral(3);
        fft2_r2c_many(mx, my, 3, pf, sf);
        for (const Mode& m : active) {
            const size_t i = m.index;
            const cplx a = sf[i], b = sf[spec + i], c = sf[2 * spec + i];
            ox[i] += -I * (m.kx * a + m.ky * b);
            oy[i] += -I * (m.kx * b + m.ky * c);
            if (rec && req_mom) {
                f_pi(0)[i] += a;
                f_pi(1)[i] += b;
                f_pi(2)[i] += c;
            }
        }
    }

    // sine forces (fcalculator.cc:702-774)
    void sine_source(const shhd_sine_force& f, std::vector<std::pair<size_t, cplx>>& out) const {
        const double grid = static_cast<double>(mx) * static_cast<double>(my);
        const cplx pos(0.0, -0.5 * f.amplitude * grid);
        const int gxp = (f.axis == 0) ? f.nk : 0;
        const int gyp = (f.axis == 0) ? 0 : f.nk;
        const int gyn = (my - f.nk) % my;
        for (const Mode& m : active) {
            cplx src(0, 0);
            if (m.gx == gxp && m.gy == gyp) src += pos;
            if (f.axis == 1 && m.gx == 0 && m.gy == gyn) src += std::conj(pos);
            if (src == cplx(0, 0)) continue;
            out.push_back({m.index, src});
        }
    }
    void add_momentum_sine_force(cplx* ox, cplx* oy) {
        for (int i = 0; i < cfg.num_sine_forces; ++i) {
            const shhd_sine_force& f = cfg.sine_forces[i];
            if (!f.momentum) continue;
            std::vector<std::pair<size_t, cplx>> src;
            sine_source(f, src);
            for (auto& p : src) (f.component == 0 ? ox : oy)[p.first] += p.second;
        }
    }
    void add_psi_sine_force(int a, cplx* out) {
        for (int i = 0; i < cfg.num_sine_forces; ++i) {
            const shhd_sine_force& f = cfg.sine_forces[i];
            if (f.momentum || f.component != a) continue;
            std::vector<std::pair<size_t, cplx>> src;
            sine_source(f, src);
            for (auto& p : src) out[p.first] += p.second;
        }
    }
    // gradient force (fcalculator.cc:776-840)
    void add_psi_gradient_force(int a, const cplx* s, cplx* out, double t) {
        bool has = false;
        for (int i = 0; i < cfg.num_gradient_forces; ++i) has = has || cfg.gradient_forces[i].component == a;
        if (!has) return;
        if (dyn == Dyn::Quiescent) throw std::runtime_error("force/gradient cannot be used in quiescent mode.");
        if (dyn == Dyn::Incompressible) {
            const double r0 = reference_density(s);
            for (int i = 0; i < cfg.num_gradient_forces; ++i) {
                const shhd_gradient_force& f = cfg.gradient_forces[i];
                if (f.component != a) continue;
                const cplx* mom = field(s, f.direction == 0 ? jx_field() : jy_field());
                for (const Mode& m : active) out[m.index] += -f.amplitude * mom[m.index] / r0;
            }
            return;
        }
        ensure_physical_fields(s, t);
        double* src = temporary_physical(1);
        for (int i = 0; i < cfg.num_gradient_forces; ++i) {
            const shhd_gradient_force& f = cfg.gradient_forces[i];
            if (f.component != a) continue;
            const double* v = phys_field(f.direction == 0 ? off_vx : off_vy);
            for (size_t p = 0; p < phys; ++p) src[p] += -f.amplitude * v[p];
        }
        cplx* sh = temporary_spectral(1);
        fft2_r2c_many(mx, my, 1, src, sh);
        for (const Mode& m : active) out[m.index] += sh[m.index];
    }

    // ---------------------------------------------------------------------
    // FCalculator stochastic parts (src/fcalculator.cc:496-700)
    // ---------------------------------------------------------------------
    void psi_sto(int a, cplx* out) {
        std::fill(out, out + spec, cplx(0, 0));
        const double mob = cfg.mobility[a];
        const double kbt = cfg.kBT * cfg.order_parameter_noise_chi;
        if (mob == 0.0 || kbt == 0.0) return;
        const double grid = static_cast<double>(mx) * static_cast<double>(my);
        const double vol = cfg.lx * cfg.ly;
        const double pref = 2.0 * kbt * grid * grid / vol;
        for (int gy = 0; gy < my; ++gy) {   // kx = 0 line, first half of each pair
            if (!is_active(0, gy)) continue;
            const int cgy = (my - gy) % my;
            if (gy >= cgy) continue;
            const double ky = ky_of(gy);
            const double k2 = ky * ky;
            if (k2 == 0.0) continue;
            const double sigma = std::sqrt(0.5 * pref * mob * k2);
            const double g1 = normal(rng);
            const double g2 = normal(rng);
            const cplx xi(g1, g2);
            const cplx fx(0.0, 0.0);
            const cplx fy = cplx(0.0, ky / k2) * sigma * xi;
            const cplx value = -I * (0.0 * fx + ky * fy);
            const size_t i = static_cast<size_t>(gy) * hx;
            const size_t ic = static_cast<size_t>(cgy) * hx;
            out[i] += value;
            out[ic] += std::conj(value);
            if (rec && req_op) {
                f_op(a, 0)[i] += fx;
                f_op(a, 1)[i] += fy;
                f_op(a, 0)[ic] += std::conj(fx);
                f_op(a, 1)[ic] += std::conj(fy);
            }
        }
        for (const Mode& m : active) {
            if (m.gx == 0 || m.k2 == 0.0) continue;
            const double sigma = std::sqrt(0.5 * pref * mob * m.k2);
            const double g1 = normal(rng);
            const double g2 = normal(rng);
            const cplx xi(g1, g2);
            const cplx fx = cplx(0.0, m.kx / m.k2) * sigma * xi;
            const cplx fy = cplx(0.0, m.ky / m.k2) * sigma * xi;
            out[m.index] += -I * (m.kx * fx + m.ky * fy);
            if (rec && req_op) {
                f_op(a, 0)[m.index] += fx;
                f_op(a, 1)[m.index] += fy;
            }
        }
    }

    void j_sto(cplx* ox, cplx* oy) {
        std::fill(ox, ox + spec, cplx(0, 0));
        std::fill(oy, oy + spec, cplx(0, 0));
        const double eta = cfg.eta, zeta = cfg.zeta, kbt = cfg.kBT;
        if (kbt == 0.0 || (eta == 0.0 && zeta == 0.0)) return;
        const double grid = static_cast<double>(mx) * static_cast<double>(my);
        const double vol = cfg.lx * cfg.ly;
        const double pref = 2.0 * kbt * grid * grid / vol;
        const double ss = std::sqrt(0.5 * pref);
        const double se = std::sqrt(eta), sz = std::sqrt(zeta);
        auto draw = [&]() { const double r = normal(rng); const double i = normal(rng); return cplx(r, i); };
        for (int gy = 0; gy < my; ++gy) {
            if (!is_active(0, gy)) continue;
            const int cgy = (my - gy) % my;
            if (gy >= cgy) continue;
            const double kx = 0.0;
            const double ky = ky_of(gy);
            if (ky * ky == 0.0) continue;
            const cplx x1 = draw(), x2 = draw(), x3 = draw();
            const cplx pxx = -ss * (sz * x1 + se * x2);
            const cplx pyy = -ss * (sz * x1 - se * x2);
            const cplx pxy = -ss * se * x3;
            const cplx vx = -I * (kx * pxx + ky * pxy);
            const cplx vy = -I * (kx * pxy + ky * pyy);
            const size_t i = static_cast<size_t>(gy) * hx;
            const size_t ic = static_cast<size_t>(cgy) * hx;
            ox[i] += vx; oy[i] += vy;
            ox[ic] += std::conj(vx); oy[ic] += std::conj(vy);
            if (rec && req_mom) {
                f_pi(0)[i] += pxx; f_pi(1)[i] += pxy; f_pi(2)[i] += pyy;
                f_pi(0)[ic] += std::conj(pxx); f_pi(1)[ic] += std::conj(pxy); f_pi(2)[ic] += std::conj(pyy);
            }
        }
        for (const Mode& m : active) {
            if (m.gx == 0 || m.k2 == 0.0) continue;
            const cplx x1 = draw(), x2 = draw(), x3 = draw();
            const cplx pxx = -ss * (sz * x1 + se * x2);
            const cplx pyy = -ss * (sz * x1 - se * x2);
            const cplx pxy = -ss * se * x3;
            ox[m.index] += -I * (m.kx * pxx + m.ky * pxy);
            oy[m.index] += -I * (m.kx * pxy + m.ky * pyy);
            if (rec && req_mom) {
                f_pi(0)[m.index] += pxx;
                f_pi(1)[m.index] += pxy;
                f_pi(2)[m.index] += pyy;
            }
        }
    }

    // ---------------------------------------------------------------------
    // initial conditions (src/initial_condition_*.cc)
    // ---------------------------------------------------------------------
    int signed_ky(int gy) const { return gy <= my / 2 ? gy : gy - my; }

    double state_rho0() const { return state[0].real() / (static_cast<double>(mx) * static_cast<double>(my)); }

    // velocity of the random-vorticity / spectrum styles before normalization
    void ic_velocity(const shhd_initial_condition& ic, const Mode& m, cplx& ux, cplx& uy) const {
        ux = uy = cplx(0, 0);
        if (m.k2 == 0.0) return;
        const int ky = signed_ky(m.gy);
        const int phase_ky = m.gx == 0 ? std::abs(ky) : ky;
        const double n = std::sqrt(static_cast<double>(m.gx) * m.gx + static_cast<double>(ky) * ky);
        cplx ph = ic_phase(ic.seed, m.gx, phase_ky);
        if (m.gx == 0 && ky < 0) ph = std::conj(ph);
        if (ic.kind == SHHD_IC_RANDOM_VORTICITY) {
            // src/initial_condition_random_vorticity_momentum.cc:56-99
            const double z = (n - ic.n0) / ic.sigma;
            const cplx omega = std::exp(-0.5 * z * z) * ph;
            ux = cplx(0.0, m.ky / m.k2) * omega;
            uy = -cplx(0.0, m.kx / m.k2) * omega;
        } else {
            // i (-ky, kx)/|k| times n^(3/2) exp(-(n/n0)^2): transverse, and
            // Hermitian on the kx = 0 line (the factor i makes the odd
            // component real in physical space)
            const double a = std::pow(n, 1.5) * std::exp(-(n / ic.n0) * (n / ic.n0));
            const double k = std::sqrt(m.k2);
            ux = I * ((-m.ky / k) * a * ph);
            uy = I * ((m.kx / k) * a * ph);
        }
    }

    void apply_ic(const shhd_initial_condition& ic) {
        const double grid = static_cast<double>(mx) * static_cast<double>(my);
        const double cell = cfg.lx * cfg.ly / grid;
        int f = -1;
        switch (ic.kind) {
            case SHHD_IC_EQUILIBRIUM_DENSITY: f = 0; break;
            case SHHD_IC_EQUILIBRIUM_ORDER_PARAMETER:
                if (ic.component < 0 || ic.component >= N) throw std::runtime_error("order-parameter index out of range");
                f = psi_field(ic.component);
                break;
            default:
                if (ic.component != 0 && ic.component != 1) throw std::runtime_error("momentum direction must be 0 or 1");
                f = jx_field() + ic.component;
        }
        cplx* out = state.data() + static_cast<size_t>(f) * spec;
        if (ic.kind == SHHD_IC_EQUILIBRIUM_DENSITY) {   // :69-93
            std::fill(out, out + spec, cplx(0, 0));
            out[0] = cplx(ic.value * grid, 0.0);
            if (ic.kBT == 0.0) return;
            const double sigma = std::sqrt(ic.kBT * grid * ic.value / (2.0 * ic.pressure_coefficient * cell));
            for (const Mode& m : active)
                if (m.k2 != 0.0) out[m.index] = sigma * ic_gaussian(ic.seed, 0, m.gx, signed_ky(m.gy));
        } else if (ic.kind == SHHD_IC_EQUILIBRIUM_ORDER_PARAMETER) {   // :73-108
            std::fill(out, out + spec, cplx(0, 0));
            out[0] = cplx(ic.value * grid, 0.0);
            if (ic.kBT == 0.0) return;
            for (const Mode& m : active) {
                if (m.k2 == 0.0) continue;
                const double coef = ic.k0 + ic.k2 * m.k2 + ic.k4 * m.k2 * m.k2;
                if (coef <= 0.0)
                    throw std::runtime_error("set order_parameter equilibrium/gaussian requires positive linear "
                                             "chemical-potential coefficient for every active nonzero mode.");
                const double sigma = std::sqrt(ic.kBT * grid / (2.0 * coef * cell));
                out[m.index] = sigma * ic_gaussian(ic.seed, ic.component, m.gx, signed_ky(m.gy));
            }
        } else if (ic.kind == SHHD_IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE ||
                   ic.kind == SHHD_IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE) {
            // compressible :90-112, incompressible :90-115
            const bool comp = ic.kind == SHHD_IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE;
            std::fill(out, out + spec, cplx(0, 0));
            if (ic.kBT == 0.0) return;
            const double rho0 = state_rho0();
            if (rho0 <= 0.0)
                throw std::runtime_error(comp ? "set momentum equilibrium/gaussian/compressible requires positive reference density."
                                              : "set momentum equilibrium/gaussian/incompressible requires positive reference density.");
            const double sigma = std::sqrt(ic.kBT * grid * rho0 / (2.0 * cell));
            for (const Mode& m : active) {
                if (m.k2 == 0.0) continue;
                if (comp) {
                    out[m.index] = sigma * ic_gaussian(ic.seed, ic.component, m.gx, signed_ky(m.gy));
                } else {
                    const double k = std::sqrt(m.k2);
                    const double c = ic.component == 0 ? -m.ky / k : m.kx / k;
                    out[m.index] = c * (sigma * ic_gaussian(ic.seed, 0, m.gx, signed_ky(m.gy)));
                }
            }
        } else {
            // random_vorticity :101-147; the spectrum style normalizes the same way
            double norm = 0.0;
            for (const Mode& m : active) {
                cplx ux, uy;
                ic_velocity(ic, m, ux, uy);
                norm += (m.gx == 0 ? 1.0 : 2.0) * (std::norm(ux) + std::norm(uy));
            }
            if (norm == 0.0) throw std::runtime_error("set momentum random_vorticity generated zero velocity field.");
            const double rho0 = state_rho0();
            if (rho0 == 0.0) throw std::runtime_error("set momentum random_vorticity requires nonzero reference density.");
            const double scale = ic.urms * grid / std::sqrt(norm);
            std::fill(out, out + spec, cplx(0, 0));
            for (const Mode& m : active) {
                cplx ux, uy;
                ic_velocity(ic, m, ux, uy);
                out[m.index] = rho0 * scale * (ic.component == 0 ? ux : uy);
            }
        }
    }

    // ---------------------------------------------------------------------
    // FluxBuffer accumulation (src/buffer_flux.cc:170-225)
    // ---------------------------------------------------------------------
    void flux_begin(std::vector<cplx>& b) {
        if (flux_requested()) std::fill(b.begin(), b.end(), cplx(0, 0));
    }
    void flux_accumulate(double w, const std::vector<cplx>& src) {
        if (!flux_requested() || w == 0.0) return;
        auto acc = [&](int k) {
            cplx* d = flux_total.data() + static_cast<size_t>(k) * spec;
            const cplx* v = src.data() + static_cast<size_t>(k) * spec;
            for (size_t i = 0; i < spec; ++i) d[i] += w * v[i];
        };
        if (req_density) { acc(0); acc(1); }
        if (req_op) for (int a = 0; a < N; ++a) { acc(2 + 2 * a); acc(3 + 2 * a); }
        if (req_mom) { acc(2 + 2 * N); acc(3 + 2 * N); acc(4 + 2 * N); }
    }
    void flux_projection_pressure() {
        if (!req_mom) return;
        cplx* pxx = flux_total.data() + static_cast<size_t>(2 + 2 * N) * spec;
        cplx* pxy = pxx + spec;
        cplx* pyy = pxy + spec;
        for (const Mode& m : active) {
            if (m.k2 == 0.0) continue;
            const size_t i = m.index;
            const cplx rx = -I * (m.kx * pxx[i] + m.ky * pxy[i]);
            const cplx ry = -I * (m.kx * pxy[i] + m.ky * pyy[i]);
            const cplx kr = m.kx * rx + m.ky * ry;
            const cplx p = -I * kr / m.k2;
            pxx[i] += p;
            pyy[i] += p;
        }
    }

    // ---------------------------------------------------------------------
    // TimeIntegrator helpers (src/time_integrator.h:53-109)
    // ---------------------------------------------------------------------
    void enforce_real_symmetry(std::vector<cplx>& s) {
        for (int f = 0; f < nfields; ++f) {
            cplx* d = field(s, f);
            for (size_t idx : inactive) d[idx] = cplx(0, 0);
        }
        for (int f = 0; f < nfields; ++f) {
            cplx* d = field(s, f);
            for (int gy = 0; gy < my; ++gy) {
                if (!is_active(0, gy)) continue;
                const int cgy = (my - gy) % my;
                if (cgy == gy) {
                    cplx& v = d[static_cast<size_t>(gy) * hx];
                    v = cplx(v.real(), 0.0);
                    continue;
                }
                if (gy > cgy) continue;
                cplx& a = d[static_cast<size_t>(gy) * hx];
                cplx& b = d[static_cast<size_t>(cgy) * hx];
                const cplx p = 0.5 * (a + std::conj(b));
                a = p;
                b = std::conj(p);
            }
        }
    }

    void eval_det(const std::vector<cplx>& cur, double t, bool with_rho) {
        std::fill(det.begin(), det.end(), cplx(0, 0));
        const cplx* c = cur.data();
        if (with_rho) rho_det(c, field(det, 0));
        for (int a = 0; a < N; ++a) psi_det(a, c, field(det, psi_field(a)), t);
        if (dyn != Dyn::Quiescent) j_det(c, field(det, jx_field()), field(det, jy_field()), t);
    }

    void draw_noise(std::vector<cplx>& dst) {
        std::fill(dst.begin(), dst.end(), cplx(0, 0));
        if (has_psi_sto) for (int a = 0; a < N; ++a) psi_sto(a, field(dst, psi_field(a)));
        if (has_j_sto) j_sto(field(dst, jx_field()), field(dst, jy_field()));
    }

    // ---------------------------------------------------------------------
    // SRK3 stage (src/time_integrator_srk3_{compressible,incompressible,quiescent}.h)
    // ---------------------------------------------------------------------
    void stage(std::vector<cplx>& next, const std::vector<cplx>& cur, const std::vector<cplx>& base,
               double t, double beta, double wb, double wc, double flux_weight) {
        const bool compressible = dyn == Dyn::Compressible;
        if (!compressible && &next != &cur) next = cur;   // copy_state(next, current)
        flux_begin(flux_det);
        rec = flux_requested() ? &flux_det : nullptr;
        eval_det(cur, t, compressible);
        rec = nullptr;
        const cplx* cd = cur.data();
        const cplx* bd = base.data();
        const cplx* dd = det.data();
        const cplx* sa = sto_a.data();
        const cplx* sb = sto_b.data();
        cplx* nd = next.data();
        if (compressible) {
            for (int f = 0; f < nfields; ++f) {
                const size_t off = static_cast<size_t>(f) * spec;
                for (const Mode& m : active) {
                    const size_t i = off + m.index;
                    const cplx st = sa[i] + beta * sb[i];
                    nd[i] = wb * bd[i] + wc * (cd[i] + dt * dd[i] + sqrt_dt * st);
                }
            }
        } else {
            for (int a = 0; a < N; ++a) {
                const size_t off = static_cast<size_t>(psi_field(a)) * spec;
                for (const Mode& m : active) {
                    const size_t i = off + m.index;
                    const cplx st = sa[i] + beta * sb[i];
                    nd[i] = wb * bd[i] + wc * (cd[i] + dt * dd[i] + sqrt_dt * st);
                }
            }
            if (dyn == Dyn::Incompressible) {
                const size_t ox = static_cast<size_t>(jx_field()) * spec;
                const size_t oy = static_cast<size_t>(jy_field()) * spec;
                for (const Mode& m : active) {
                    const size_t i = m.index;
                    const cplx sx = sa[ox + i] + beta * sb[ox + i];
                    const cplx sy = sa[oy + i] + beta * sb[oy + i];
                    jinc_x[i] = dt * dd[ox + i] + sqrt_dt * sx;
                    jinc_y[i] = dt * dd[oy + i] + sqrt_dt * sy;
                    if (m.k2 != 0.0) {
                        const cplx tr = -m.ky * jinc_x[i] + m.kx * jinc_y[i];
                        jinc_x[i] = -m.ky * tr / m.k2;
                        jinc_y[i] = m.kx * tr / m.k2;
                    }
                }
                for (const Mode& m : active) {
                    const size_t i = m.index;
                    nd[ox + i] = wb * bd[ox + i] + wc * (cd[ox + i] + jinc_x[i]);
                    nd[oy + i] = wb * bd[oy + i] + wc * (cd[oy + i] + jinc_y[i]);
                }
            }
        }
        enforce_real_symmetry(next);
        flux_accumulate(flux_weight, flux_det);
    }

    void step_srk3(double t) {
        static constexpr double S2 = 1.41421356237309504880;
        static constexpr double S3 = 1.73205080756887729352;
        static constexpr double B1 = (2.0 * S2 + S3) / 5.0;
        static constexpr double B2 = (-4.0 * S2 + 3.0 * S3) / 5.0;
        static constexpr double B3 = (S2 - 2.0 * S3) / 10.0;
        u_old = state;
        flux_begin(flux_total);
        flux_begin(flux_sto);
        std::fill(sto_a.begin(), sto_a.end(), cplx(0, 0));
        std::fill(sto_b.begin(), sto_b.end(), cplx(0, 0));
        // draw order: psi A/B per order parameter, then j A, j B
        // (time_integrator_srk3_compressible.h:116-126)
        std::vector<cplx>* sto_rec = flux_requested() ? &flux_sto : nullptr;
        if (has_psi_sto) {
            for (int a = 0; a < N; ++a) {
                rec = sto_rec;
                psi_sto(a, field(sto_a, psi_field(a)));
                rec = nullptr;
                psi_sto(a, field(sto_b, psi_field(a)));
            }
        }
        if (has_j_sto) {
            rec = sto_rec;
            j_sto(field(sto_a, jx_field()), field(sto_a, jy_field()));
            rec = nullptr;
            j_sto(field(sto_b, jx_field()), field(sto_b, jy_field()));
        }
        stage(u1, u_old, u_old, t, B1, 0.0, 1.0, 1.0 / 6.0);
        stage(u2, u1, u_old, t + dt, B2, 3.0 / 4.0, 1.0 / 4.0, 1.0 / 6.0);
        stage(state, u2, u_old, t + 0.5 * dt, B3, 1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0);
        flux_accumulate(1.0 / sqrt_dt, flux_sto);
        if (dyn == Dyn::Incompressible) flux_projection_pressure();
    }

    // Euler (src/time_integrator_euler_*.h): one draw, interleaved with det
    // per order parameter exactly as the reference loops.
    void step_euler(double t) {
        u_old = state;
        flux_begin(flux_total);
        flux_begin(flux_det);
        flux_begin(flux_sto);
        std::vector<cplx>* det_rec = flux_requested() ? &flux_det : nullptr;
        std::vector<cplx>* sto_rec = flux_requested() ? &flux_sto : nullptr;
        std::fill(det.begin(), det.end(), cplx(0, 0));
        std::fill(sto_a.begin(), sto_a.end(), cplx(0, 0));
        const cplx* c = u_old.data();
        const bool compressible = dyn == Dyn::Compressible;
        rec = det_rec;
        if (compressible) rho_det(c, field(det, 0));
        for (int a = 0; a < N; ++a) {
            rec = det_rec;
            psi_det(a, c, field(det, psi_field(a)), t);
            rec = sto_rec;
            if (has_psi_sto) psi_sto(a, field(sto_a, psi_field(a)));
        }
        if (dyn != Dyn::Quiescent) {
            rec = det_rec;
            j_det(c, field(det, jx_field()), field(det, jy_field()), t);
            rec = sto_rec;
            if (has_j_sto) j_sto(field(sto_a, jx_field()), field(sto_a, jy_field()));
        }
        rec = nullptr;
        cplx* nd = state.data();
        const cplx* dd = det.data();
        const cplx* sd = sto_a.data();
        if (compressible) {
            for (int f = 0; f < nfields; ++f) {
                const size_t off = static_cast<size_t>(f) * spec;
                for (const Mode& m : active) {
                    const size_t i = off + m.index;
                    nd[i] = c[i] + dt * dd[i] + sqrt_dt * sd[i];
                }
            }
        } else {
            for (int a = 0; a < N; ++a) {
                const size_t off = static_cast<size_t>(psi_field(a)) * spec;
                for (const Mode& m : active) {
                    const size_t i = off + m.index;
                    nd[i] = c[i] + dt * dd[i] + sqrt_dt * sd[i];
                }
            }
            if (dyn == Dyn::Incompressible) {
                const size_t ox = static_cast<size_t>(jx_field()) * spec;
                const size_t oy = static_cast<size_t>(jy_field()) * spec;
                for (const Mode& m : active) {
                    const size_t i = m.index;
                    jinc_x[i] = dt * dd[ox + i] + sqrt_dt * sd[ox + i];
                    jinc_y[i] = dt * dd[oy + i] + sqrt_dt * sd[oy + i];
                    if (m.k2 != 0.0) {
                        const cplx tr = -m.ky * jinc_x[i] + m.kx * jinc_y[i];
                        jinc_x[i] = -m.ky * tr / m.k2;
                        jinc_y[i] = m.kx * tr / m.k2;
                    }
                }
                for (const Mode& m : active) {
                    const size_t i = m.index;
                    nd[ox + i] = c[ox + i] + jinc_x[i];
                    nd[oy + i] = c[oy + i] + jinc_y[i];
                }
            }
        }
        enforce_real_symmetry(state);
        flux_accumulate(1.0, flux_det);
        flux_accumulate(1.0 / sqrt_dt, flux_sto);
        if (dyn == Dyn::Incompressible) flux_projection_pressure();
    }
};

std::string g_create_error;

int fail(Oracle* o, const std::exception& e) {
    if (o) o->error = e.what();
    return SHHD_ERR_RUNTIME;
}

}  // namespace

struct oracle_solver {
    Oracle* impl;
};

extern "C" {

int oracle_create(const shhd_config* cfg, oracle_solver** out) {
    if (!cfg || !out) return SHHD_ERR_INVALID;
    try {
        auto* s = new oracle_solver;
        s->impl = new Oracle(*cfg);
        *out = s;
        g_create_error.clear();
        return SHHD_OK;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        *out = nullptr;
        return SHHD_ERR_INVALID;
    }
}

const char* oracle_create_error(void) { return g_create_error.c_str(); }

void oracle_destroy(oracle_solver* s) {
    if (!s) return;
    delete s->impl;
    delete s;
}

const char* oracle_last_error(const oracle_solver* s) { return s ? s->impl->error.c_str() : g_create_error.c_str(); }

int oracle_set_state(oracle_solver* s, const double* in, size_t n) {
    Oracle& o = *s->impl;
    if (n != 2 * o.state.size()) { o.error = "state size mismatch"; return SHHD_ERR_INVALID; }
    std::memcpy(o.state.data(), in, sizeof(double) * n);
    o.physical_ready = false;
    o.cached_state = nullptr;
    o.rho0_ready = false;   // a new state starts a new run
    return SHHD_OK;
}

int oracle_get_state(const oracle_solver* s, double* out, size_t n) {
    const Oracle& o = *s->impl;
    if (n != 2 * o.state.size()) return SHHD_ERR_INVALID;
    std::memcpy(out, o.state.data(), sizeof(double) * n);
    return SHHD_OK;
}

int oracle_step(oracle_solver* s, int nsteps, double t0) {
    Oracle& o = *s->impl;
    try {
        for (int i = 0; i < nsteps; ++i) {
            const double t = t0 + static_cast<double>(i) * o.dt;
            if (o.srk3) o.step_srk3(t); else o.step_euler(t);
        }
        return SHHD_OK;
    } catch (const std::exception& e) {
        return fail(&o, e);
    }
}

int oracle_rhs_det(oracle_solver* s, const double* state, double* out, double t) {
    Oracle& o = *s->impl;
    try {
        std::vector<cplx> cur(o.state.size());
        std::memcpy(cur.data(), state, sizeof(cplx) * cur.size());
        o.physical_ready = false;
        o.cached_state = nullptr;
        o.eval_det(cur, t, o.dyn == Dyn::Compressible);
        o.physical_ready = false;
        o.cached_state = nullptr;
        std::memcpy(out, o.det.data(), sizeof(cplx) * o.det.size());
        return SHHD_OK;
    } catch (const std::exception& e) {
        return fail(&o, e);
    }
}

int oracle_rhs_sto(oracle_solver* s, double* out) {
    Oracle& o = *s->impl;
    o.draw_noise(o.sto_a);
    std::memcpy(out, o.sto_a.data(), sizeof(cplx) * o.sto_a.size());
    return SHHD_OK;
}

int oracle_set_flux_request(oracle_solver* s, int density, int order_parameter, int momentum) {
    Oracle& o = *s->impl;
    o.req_density = density != 0;
    o.req_op = order_parameter != 0 && o.N > 0;
    o.req_mom = momentum != 0;
    return SHHD_OK;
}

int oracle_get_flux(const oracle_solver* s, double* out, size_t n) {
    const Oracle& o = *s->impl;
    if (n != 2 * o.flux_total.size()) return SHHD_ERR_INVALID;
    std::memcpy(out, o.flux_total.data(), sizeof(double) * n);
    return SHHD_OK;
}

int oracle_apply_initial_condition(oracle_solver* s, const shhd_initial_condition* ic) {
    Oracle& o = *s->impl;
    try {
        o.apply_ic(*ic);
        o.physical_ready = false;
        o.cached_state = nullptr;
        o.rho0_ready = false;
        return SHHD_OK;
    } catch (const std::exception& e) {
        o.error = e.what();
        return SHHD_ERR_INVALID;
    }
}

unsigned long long oracle_splitmix64(unsigned long long v) { return ic_splitmix64(v); }

int oracle_fft_c2c(int n, const double* in, double* out, int sign) {
    try {
        auto p = plan_for(n);
        std::vector<cplx> d(static_cast<size_t>(n)), scratch;
        std::memcpy(d.data(), in, sizeof(cplx) * n);
        fft_inplace(*p, d.data(), scratch, sign < 0 ? -1 : 1);
        std::memcpy(out, d.data(), sizeof(cplx) * n);
        return SHHD_OK;
    } catch (...) {
        return SHHD_ERR_INVALID;
    }
}

int oracle_fft2_r2c(int mx, int my, const double* phys, double* spec) {
    fft2_r2c_many(mx, my, 1, phys, reinterpret_cast<cplx*>(spec));
    return SHHD_OK;
}

int oracle_fft2_c2r(int mx, int my, const double* spec, double* phys) {
    std::vector<cplx> work;
    fft2_c2r_many(mx, my, 1, reinterpret_cast<const cplx*>(spec), phys, work);
    return SHHD_OK;
}

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // extern "C"
