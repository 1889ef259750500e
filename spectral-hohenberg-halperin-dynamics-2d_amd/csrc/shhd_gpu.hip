// shhd_gpu.hip — C-ABI of the MI355X right-hand-side and time-step engine
// (declared in include/shhd_gpu.h).
//
// One handle = one simulation on one device stream.  The state lives in HBM in
// a compact spectral layout holding active modes only:
//     field f, column gx < Nax/2, slot c < Nay  ->  index f*(Nax/2*Nay) + gx*Nay + c
// with ky = c for c < Nay/2 and ky = c - Nay otherwise (slot Nay/2 is the
// always-zero ky = -Nay/2 mode).  Per stage the engine runs
//     K1 (inverse y, F_in fields) -> K2 (rows: x transforms + products)
//     -> K3 (forward y + spectral RHS + noise + stage update),
// so one SRK3 step is 9 launches and one Euler step 3.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/shhd_gpu.h"
#include "instances.h"

using namespace shhd;

namespace {

std::string g_create_error;

constexpr int kClasses = 3;
constexpr size_t kLdsCap = 96 * 1024;
constexpr size_t kLdsCap3 = 128 * 1024;

struct ColPlan {
    int C = 1, NT = 256, V = 6, ntiles = 1;
    int acc = 6;       // K3 accumulator slots per lane: 4 when C*Nay <= 4*NT, else V
    size_t lds = 0;    // K1: C*My values
    size_t lds3 = 0;   // K3: max(C*My, 2*C*Nay) values (FFT column, then both accumulators)
};
struct RowPlan {
    int R = 1, NT = 256, V = 6, nblocks = 1, maxn = 0;
    size_t lds = 0;
};

// device transform lengths: 2^a or 3 * 2^a (every Stockham sub-length ns is
// then a power of two, which the pass indexing relies on)
bool is_device_length(int n) {
    if (n <= 0) return false;
    if (n % 3 == 0) n /= 3;
    while (n % 2 == 0) n /= 2;
    return n == 1;
}

FftPlanDev make_fft_plan(int L) {
    FftPlanDev p{};
    std::vector<int> rad;
    int m = L;
    while (m % 8 == 0 && m > 0) { rad.push_back(8); m /= 8; }
    while (m % 4 == 0) { rad.push_back(4); m /= 4; }
    while (m % 2 == 0) { rad.push_back(2); m /= 2; }
    while (m % 3 == 0) { rad.push_back(3); m /= 3; }   // radix 3 last: every ns stays a power of two
    int ns = 1;
    for (int r : rad) {
        FftPass& ps = p.pass[p.npass++];
        ps.radix = r;
        ps.m = L / r;
        int lns = 0;
        while ((1 << lns) < ns) ++lns;
        ps.lns = lns;
        ps.tstep = L / (ns * r);
        ps.inv_m = 1.0f / static_cast<float>(ps.m);
        ns *= r;
    }
    return p;
}

int pow2ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

ColPlan plan_columns(int my, int nay, int hxa, size_t elem) {
    ColPlan p;
    p.V = (my % 3 == 0) ? 6 : 8;
    int target = static_cast<int>(32 / elem);   // 32-byte row segments (measured best at 1536^2)
    if (const char* e = std::getenv("SHHD_COLUMN_TILE")) target = std::max(1, std::atoi(e));   // tuning override
    int C = std::min(std::max(target, 1), pow2ceil(hxa));
    auto lds3 = [&](int c) { return static_cast<size_t>(std::max(lp_size(c * my), 2 * c * nay)) * elem; };
    while (C > 1 && (lds3(C) > kLdsCap3 || C * my > 1024 * p.V)) C >>= 1;
    p.C = C;
    const int n = C * my;
    p.NT = (n <= 64 * p.V) ? 64 : (n <= 256 * p.V ? 256 : (n <= 512 * p.V ? 512 : 1024));
    p.ntiles = (hxa + C - 1) / C;
    p.acc = (C * nay <= 4 * p.NT) ? 4 : p.V;
    p.lds = static_cast<size_t>(lp_size(C * my)) * elem;
    p.lds3 = lds3(C);
    return p;
}

RowPlan plan_rows(int mx, int my, size_t elem) {
    RowPlan p;
    p.V = (mx % 3 == 0) ? 6 : 8;
    p.NT = (mx <= 64 * p.V) ? 64 : (mx <= 256 * p.V ? 256 : (mx <= 512 * p.V ? 512 : 1024));
    p.R = std::max(1, (p.NT * p.V) / mx);
    while (p.R > 1 && static_cast<size_t>(lp_size(p.R * mx)) * elem > kLdsCap) --p.R;
    p.R = std::min(p.R, my);
    p.nblocks = (my + p.R - 1) / p.R;
    p.lds = static_cast<size_t>(lp_size(p.R * mx)) * elem;
    return p;
}

// ---------------------------------------------------------------------------
// kernel dispatch over the instantiated (NT, V, MAXN) shapes
// ---------------------------------------------------------------------------
template <typename T>
using K1Fn = void (*)(Desc, StagePtrs<T>, cx<T>*, const cx<T>*, int);
template <typename T>
using K2Fn = void (*)(Desc, const cx<T>*, cx<T>*, const cx<T>*, int, int*);
template <typename T>
using K3Fn = void (*)(Desc, StagePtrs<T>, StageParams, const cx<T>*, const cx<T>*, int);

template <typename T>
K1Fn<T> pick_k1(int NT, int V) {
#define SHHD_K1(nt, v) if (NT == nt && V == v) return &k_inverse_columns<T, nt, v>;
    SHHD_FOR_NT_V(SHHD_K1)
#undef SHHD_K1
    return nullptr;
}

template <typename T>
K3Fn<T> pick_k3(int NT, int V, int acc) {
#define SHHD_K3(nt, v) \
    if (NT == nt && V == v && acc == 4) return &k_forward_columns<T, nt, v, 4>; \
    if (NT == nt && V == v && acc == v) return &k_forward_columns<T, nt, v, v>;
    SHHD_FOR_NT_V(SHHD_K3)
#undef SHHD_K3
    return nullptr;
}

template <typename T>
K2Fn<T> pick_k2(int NT, int V, int maxn) {
    const int n = maxn <= 0 ? 0 : (maxn == 1 ? 1 : (maxn == 2 ? 2 : (maxn <= 4 ? 4 : 8)));
#define SHHD_K2N(nt, v, mn) if (NT == nt && V == v && n == mn) return &k_rows<T, nt, v, mn>;
#define SHHD_K2(nt, v) SHHD_K2N(nt, v, 0) SHHD_K2N(nt, v, 1) SHHD_K2N(nt, v, 2) SHHD_K2N(nt, v, 4) SHHD_K2N(nt, v, 8)
    SHHD_FOR_NT_V(SHHD_K2)
#undef SHHD_K2
#undef SHHD_K2N
    return nullptr;
}

__global__ void k_clear_flag(int* flag) { *flag = 0; }

// Hermitian projection + self-conjugate fix of the kx = 0 column of fields
// that no stage rewrites (rho when incompressible, rho and j when quiescent);
// idempotent, so it runs at the start of every step() call.
template <typename T>
__global__ void k_enforce_static(cx<T>* base, int nay, size_t field_stride, const int* fields, int nfields_static,
                                 cx<double>* k0) {
    const int f = fields[blockIdx.x];
    cx<T>* col = base + static_cast<size_t>(f) * field_stride;   // gx = 0 column
    const int half = nay / 2;
    for (int m = threadIdx.x; m < half; m += blockDim.x) {
        if (m == 0) {
            col[0].im = T(0);
            if (k0) k0[f].im = 0.0;
        } else {
            const cx<T> a = col[m];
            const cx<T> b = col[nay - m];
            const cx<T> p = T(0.5) * (a + conj(b));
            col[m] = p;
            col[nay - m] = conj(p);
        }
    }
    (void)nfields_static;
}

}  // namespace

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
struct shhd_gpu {
    shhd_config cfg{};
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    bool fp32 = false;
    size_t elem = 16;   // bytes per complex value

    Desc desc{};
    ColPlan cols;
    RowPlan rows;
    int n_in = 0, n_out = 0, n_updated = 0, stages = 3;

    size_t field_elems = 0;   // hxa * nay
    size_t plane_elems = 0;   // my * hxa
    void* state[3] = {nullptr, nullptr, nullptr};
    cx<double>* k0[3] = {nullptr, nullptr, nullptr};
    void* tbuf = nullptr;
    void* ubuf = nullptr;
    void* twx = nullptr;
    void* twy = nullptr;
    int* dflag = nullptr;
    int* dstatic = nullptr;
    std::vector<int> static_fields;
    bool is_static[kMaxFields] = {};
    int cur = 0;
    uint64_t step_index = 0;
    bool state_set = false;
    bool rho0_needed = false;

    bool timing = false;
    double time_ms[kClasses] = {0, 0, 0};
    long long launches[kClasses] = {0, 0, 0};
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;

    size_t device_bytes = 0;
};

namespace {

shhd_status set_err(shhd_gpu* h, shhd_status st, const std::string& msg) {
    if (h) h->err = msg;
    else g_create_error = msg;
    return st;
}

#define SHHD_HIP(h, call)                                                                      \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return set_err(h, SHHD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
void make_twiddles(int L, std::vector<cx<T>>& out) {
    out.resize(static_cast<size_t>(L));
    const double pi = 3.14159265358979323846;
    for (int j = 0; j < L; ++j) {
        const double a = 2.0 * pi * static_cast<double>(j) / static_cast<double>(L);
        out[static_cast<size_t>(j)] = mk(static_cast<T>(std::cos(a)), static_cast<T>(-std::sin(a)));
    }
}

// Lower shhd_config into the kernel descriptor (see desc.h for the meaning of
// each entry).  Mirrors the reference's request logic:
// PhysicalRHSFieldRequests (src/fcalculator_field_requests_and_layout.h:30-127),
// the early returns of psi_det / j_det (src/fcalculator.cc:86-150) and the
// noise enablement (src/fluid_solver.h:79-94).
shhd_status build_desc(shhd_gpu* h) {
    const shhd_config& c = h->cfg;
    Desc& d = h->desc;
    d = Desc{};
    const int N = c.num_order_parameters;
    const bool quiescent = c.mode == SHHD_MODE_QUIESCENT;
    const bool compressible = c.mode == SHHD_MODE_COMPRESSIBLE;
    const bool incompressible = c.mode == SHHD_MODE_INCOMPRESSIBLE;

    d.nax = c.nx_active;
    d.nay = c.ny_active;
    d.mx = c.nx;
    d.my = c.ny;
    d.hxa = c.nx_active / 2;
    d.lx = c.lx;
    d.ly = c.ly;
    d.N = N;
    d.nfields = N + 3;
    d.mode = c.mode;
    d.in_scale = 1.0 / (static_cast<double>(c.nx) * static_cast<double>(c.ny));
    d.dt = c.dt;
    d.sqrt_dt = std::sqrt(c.dt);

    bool fe_phys_mu = false;
    if (c.free_energy_kind == SHHD_FREE_ENERGY_PHI4)
        for (int a = 0; a < N; ++a) fe_phys_mu = fe_phys_mu || c.phi4_u[a] != 0.0;
    const bool psi_adv = c.order_parameter_advection != 0 && N > 0;
    const bool mom_adv = c.momentum_advection != 0 && !quiescent;
    const bool visc = c.eta != 0.0 || c.zeta != 0.0;

    for (int a = 0; a < N; ++a) {
        d.mobility[a] = c.mobility[a];
        d.mu_k0[a] = c.mu_k0[a];
        d.mu_k2[a] = c.mu_k2[a];
        d.mu_k4[a] = c.mu_k4[a];
        d.phi4_u[a] = c.free_energy_kind == SHHD_FREE_ENERGY_PHI4 ? c.phi4_u[a] : 0.0;
        d.psi_det_on[a] = (c.mobility[a] != 0.0 || psi_adv) ? 1 : 0;
        d.psi_lin_on[a] = (c.mobility[a] != 0.0 && (c.mu_k0[a] != 0.0 || c.mu_k2[a] != 0.0 || c.mu_k4[a] != 0.0)) ? 1 : 0;
    }
    d.mom_det_on = (!quiescent && (c.pressure_coefficient != 0.0 || visc || mom_adv)) ? 1 : 0;
    d.pressure_coefficient = c.pressure_coefficient;
    d.eta = c.eta;
    d.zeta = c.zeta;
    d.incompressible_viscous = (incompressible && c.eta != 0.0) ? 1 : 0;

    // products of the row pass
    auto add_out = [&](int kind, int alpha) {
        d.out[d.n_out].kind = kind;
        d.out[d.n_out].alpha = alpha;
        return d.n_out++;
    };
    int grad_count[kMaxOp] = {};
    for (int q = 0; q < c.num_gradient_forces; ++q) {
        const shhd_gradient_force& f = c.gradient_forces[q];
        grad_count[f.component]++;
    }
    std::vector<int> psi_contrib[kMaxOp];
    bool need_psi[kMaxOp] = {};
    bool need_vel = false, need_j = false;
    for (int a = 0; a < N; ++a) {
        if (c.mobility[a] != 0.0 && fe_phys_mu) { psi_contrib[a].push_back(add_out(OUT_MU, a)); need_psi[a] = true; }
        if (psi_adv) {
            psi_contrib[a].push_back(add_out(OUT_PSIVX, a));
            psi_contrib[a].push_back(add_out(OUT_PSIVY, a));
            need_psi[a] = true;
            need_vel = true;
        }
        if (compressible && d.psi_det_on[a] && grad_count[a] > 0) {
            psi_contrib[a].push_back(add_out(OUT_GRADF, a));
            need_vel = true;
        }
    }
    std::vector<int> mom_contrib;
    if (compressible && visc) {
        mom_contrib.push_back(add_out(OUT_VX, 0));
        mom_contrib.push_back(add_out(OUT_VY, 0));
        need_vel = true;
    }
    if (mom_adv) {
        mom_contrib.push_back(add_out(OUT_PXX, 0));
        mom_contrib.push_back(add_out(OUT_PXY, 0));
        mom_contrib.push_back(add_out(OUT_PYY, 0));
        need_vel = true;
        need_j = true;
    }
    if (need_vel) need_j = true;
    for (int q = 0; q < c.num_gradient_forces; ++q) {
        const shhd_gradient_force& f = c.gradient_forces[q];
        if (compressible) {
            d.grad_component[d.n_grad] = f.component;
            d.grad_direction[d.n_grad] = f.direction;
            d.grad_amplitude[d.n_grad] = f.amplitude;
            d.n_grad++;
        } else if (incompressible && d.psi_det_on[f.component]) {
            d.igrad_component[d.n_igrad] = f.component;
            d.igrad_direction[d.n_igrad] = f.direction;
            d.igrad_amplitude[d.n_igrad] = f.amplitude;
            d.n_igrad++;
        }
    }

    // inverse-transformed fields, in the reference's pack order
    // (PhysicalFieldLayout::assign_offsets: psi, rho, jx, jy)
    for (int a = 0; a < N; ++a) {
        if (!need_psi[a]) continue;
        d.in_field[d.n_in] = 1 + a;
        d.in_role[d.n_in] = ROLE_PSI0 + a;
        d.n_in++;
    }
    if (need_vel && compressible) {
        d.in_field[d.n_in] = 0;
        d.in_role[d.n_in] = ROLE_RHO;
        d.n_in++;
    }
    if (need_j) {
        d.in_field[d.n_in] = N + 1;
        d.in_role[d.n_in] = ROLE_JX;
        d.n_in++;
        d.in_field[d.n_in] = N + 2;
        d.in_role[d.n_in] = ROLE_JY;
        d.n_in++;
    }
    d.make_velocity = need_vel ? 1 : 0;
    d.velocity_compressible = compressible ? 1 : 0;
    h->rho0_needed = incompressible && (need_vel || d.incompressible_viscous || d.n_igrad > 0);

    // groups
    for (int a = 0; a < N; ++a) {
        Group& g = d.groups[d.ngroups++];
        g.kind = GROUP_PSI;
        g.alpha = a;
        g.ncontrib = static_cast<int>(psi_contrib[a].size());
        for (int i = 0; i < g.ncontrib; ++i) g.contrib[i] = psi_contrib[a][static_cast<size_t>(i)];
    }
    if (!quiescent) {
        Group& g = d.groups[d.ngroups++];
        g.kind = GROUP_MOM;
        g.alpha = 0;
        g.ncontrib = static_cast<int>(mom_contrib.size());
        for (int i = 0; i < g.ncontrib; ++i) g.contrib[i] = mom_contrib[static_cast<size_t>(i)];
    }

    // sine forces as constant spectral sources (src/fcalculator.cc:702-774)
    const double grid = static_cast<double>(c.nx) * static_cast<double>(c.ny);
    auto active = [&](int gx, int gy) {
        if (gx < 0 || gx >= c.nx_active / 2) return false;
        const int ky = (gy <= c.ny / 2) ? gy : gy - c.ny;
        return std::abs(ky) < c.ny_active / 2;
    };
    auto add_source = [&](int field, int gx, int gy, double re, double im) {
        if (!active(gx, gy)) return;
        const int ky = (gy <= c.ny / 2) ? gy : gy - c.ny;
        Source& s = d.sources[d.nsources++];
        s.field = field;
        s.gx = gx;
        s.cidx = ky >= 0 ? ky : ky + c.ny_active;
        s.re = re;
        s.im = im;
    };
    for (int q = 0; q < c.num_sine_forces; ++q) {
        const shhd_sine_force& f = c.sine_forces[q];
        int field;
        if (f.momentum) {
            if (!d.mom_det_on) continue;
            field = f.component == 0 ? N + 1 : N + 2;
        } else {
            if (!d.psi_det_on[f.component]) continue;
            field = 1 + f.component;
        }
        const double im = -0.5 * f.amplitude * grid;
        const int gxp = (f.axis == 0) ? f.nk : 0;
        const int gyp = (f.axis == 0) ? 0 : f.nk;
        add_source(field, gxp, gyp, 0.0, im);
        if (f.axis == 1) add_source(field, 0, (c.ny - f.nk) % c.ny, 0.0, -im);
    }

    // noise (src/fcalculator.cc:496-700)
    const double vol = c.lx * c.ly;
    const bool has_psi_sto = c.noise_order_parameter && N > 0;
    const bool has_j_sto = c.noise_momentum && !quiescent;
    const double kbt_psi = c.kBT * c.order_parameter_noise_chi;
    for (int a = 0; a < N; ++a) {
        d.psi_noise_on[a] = (has_psi_sto && c.mobility[a] != 0.0 && kbt_psi != 0.0) ? 1 : 0;
        const double pref = 2.0 * kbt_psi * grid * grid / vol;
        d.psi_noise_pref[a] = 0.5 * pref * c.mobility[a];
    }
    d.j_noise_on = (has_j_sto && c.kBT != 0.0 && visc) ? 1 : 0;
    {
        const double pref = 2.0 * c.kBT * grid * grid / vol;
        d.j_sigma = std::sqrt(0.5 * pref);
        d.j_sqrt_eta = std::sqrt(c.eta);
        d.j_sqrt_zeta = std::sqrt(c.zeta);
    }
    d.seed_lo = static_cast<uint32_t>(c.seed);
    d.seed_hi = static_cast<uint32_t>(c.seed >> 32) ^ 0x5348u;

    d.plan_x = make_fft_plan(c.nx);
    d.plan_y = make_fft_plan(c.ny);

    h->n_in = d.n_in;
    h->n_out = d.n_out;
    h->n_updated = N + (quiescent ? 0 : (compressible ? 3 : 2));
    h->static_fields.clear();
    for (int f = 0; f < N + 3; ++f) h->is_static[f] = false;
    if (!compressible) {
        h->static_fields.push_back(0);
        h->is_static[0] = true;
    }
    if (quiescent) {
        h->static_fields.push_back(N + 1);
        h->static_fields.push_back(N + 2);
        h->is_static[N + 1] = h->is_static[N + 2] = true;
    }
    return SHHD_OK;
}

shhd_status validate(const shhd_config& c) {
    const int N = c.num_order_parameters;
    if (N < 0 || N > SHHD_MAX_ORDER_PARAMETERS) return set_err(nullptr, SHHD_ERR_INVALID, "num_order_parameters out of range");
    if (c.nx_active <= 0 || c.ny_active <= 0 || c.nx_active % 2 || c.ny_active % 2)
        return set_err(nullptr, SHHD_ERR_INVALID, "grid must be even.");
    if (c.nx < c.nx_active || c.ny < c.ny_active)
        return set_err(nullptr, SHHD_ERR_INVALID, "SpectralMask2D active grid cannot exceed compute grid.");
    if (!(c.lx > 0.0) || !(c.ly > 0.0)) return set_err(nullptr, SHHD_ERR_INVALID, "domain lengths must be positive");
    if (!(c.dt > 0.0)) return set_err(nullptr, SHHD_ERR_INVALID, "timestep must be positive");
    if (c.mode < 0 || c.mode > 2) return set_err(nullptr, SHHD_ERR_INVALID, "unknown time_evolution type");
    if (c.integrator != SHHD_INTEGRATOR_EULER && c.integrator != SHHD_INTEGRATOR_SRK3)
        return set_err(nullptr, SHHD_ERR_INVALID, "unknown time_evolution type");
    if (c.precision != SHHD_PRECISION_FP64 && c.precision != SHHD_PRECISION_FP32)
        return set_err(nullptr, SHHD_ERR_INVALID, "precision must be 64 or 32");
    if (c.eta < 0.0) return set_err(nullptr, SHHD_ERR_INVALID, "transport coefficient requires nonnegative eta.");
    if (c.zeta < 0.0) return set_err(nullptr, SHHD_ERR_INVALID, "transport coefficient requires nonnegative zeta.");
    for (int a = 0; a < N; ++a) {
        if (c.mobility[a] < 0.0)
            return set_err(nullptr, SHHD_ERR_INVALID, "transport coefficient requires nonnegative order parameter mobility.");
        if (c.free_energy_kind == SHHD_FREE_ENERGY_PHI4 && c.phi4_u[a] < 0.0)
            return set_err(nullptr, SHHD_ERR_INVALID, "phi4 free energy requires nonnegative u coefficient.");
    }
    if (c.free_energy_kind < 0 || c.free_energy_kind > 2)
        return set_err(nullptr, SHHD_ERR_UNSUPPORTED, "free-energy style has no device functor");
    if (c.kBT < 0.0) return set_err(nullptr, SHHD_ERR_INVALID, "fix noise requires nonnegative kBT.");
    if (c.order_parameter_noise_chi < 0.0) return set_err(nullptr, SHHD_ERR_INVALID, "fix noise requires nonnegative chi.");
    const bool quiescent = c.mode == SHHD_MODE_QUIESCENT;
    // ParamParser::validate_configuration (src/param_parser.cc:771-805)
    if (c.order_parameter_advection && N == 0)
        return set_err(nullptr, SHHD_ERR_INVALID, "fix nonlinear order_parameter requires at least one order parameter.");
    if (quiescent && c.momentum_advection)
        return set_err(nullptr, SHHD_ERR_INVALID, "quiescent time evolution cannot use momentum nonlinear advection.");
    if (quiescent && c.order_parameter_advection)
        return set_err(nullptr, SHHD_ERR_INVALID, "quiescent time evolution cannot use order_parameter nonlinear advection.");
    if (c.num_sine_forces < 0 || c.num_sine_forces > SHHD_MAX_FORCES || c.num_gradient_forces < 0 ||
        c.num_gradient_forces > SHHD_MAX_FORCES)
        return set_err(nullptr, SHHD_ERR_INVALID, "too many forces");
    for (int q = 0; q < c.num_sine_forces; ++q) {
        const shhd_sine_force& f = c.sine_forces[q];
        if (quiescent && f.momentum)
            return set_err(nullptr, SHHD_ERR_INVALID, "quiescent time evolution cannot use momentum force/sine.");
        if (f.axis != 0 && f.axis != 1) return set_err(nullptr, SHHD_ERR_INVALID, "fix force/sine axis must be x|y|0|1.");
        if (f.nk <= 0) return set_err(nullptr, SHHD_ERR_INVALID, "fix force/sine requires nk > 0.");
        const int active = f.axis == 0 ? c.nx_active : c.ny_active;
        if (f.nk >= active / 2) return set_err(nullptr, SHHD_ERR_INVALID, "fix force/sine requires nk < active_N_axis/2.");
        if (f.momentum ? (f.component != 0 && f.component != 1) : (f.component < 0 || f.component >= N))
            return set_err(nullptr, SHHD_ERR_INVALID, "fix force/sine component out of range");
    }
    for (int q = 0; q < c.num_gradient_forces; ++q) {
        const shhd_gradient_force& f = c.gradient_forces[q];
        if (quiescent) return set_err(nullptr, SHHD_ERR_INVALID, "quiescent time evolution cannot use force/gradient.");
        if (f.component < 0 || f.component >= N)
            return set_err(nullptr, SHHD_ERR_INVALID, "fix order_parameter force/gradient: order-parameter index out of range");
        if (f.direction != 0 && f.direction != 1)
            return set_err(nullptr, SHHD_ERR_INVALID, "fix force/gradient direction must be x|y|0|1.");
    }
    // device transform shapes
    if (!is_device_length(c.nx) || !is_device_length(c.ny))
        return set_err(nullptr, SHHD_ERR_UNSUPPORTED, "compute grid sides must be 2^a or 3*2^a on the device");
    const int vx = (c.nx % 3 == 0) ? 6 : 8, vy = (c.ny % 3 == 0) ? 6 : 8;
    if (c.nx > 1024 * vx || c.ny > 1024 * vy)
        return set_err(nullptr, SHHD_ERR_UNSUPPORTED, "compute grid side exceeds the device transform limit (6144 or 8192)");
    if (c.nx < 4 || c.ny < 4) return set_err(nullptr, SHHD_ERR_UNSUPPORTED, "compute grid side must be at least 4");
    return SHHD_OK;
}

template <typename T>
StagePtrs<T> stage_ptrs(shhd_gpu* h, int next, int cur, int base) {
    StagePtrs<T> p{};
    const int nf = h->desc.nfields;
    for (int f = 0; f < nf; ++f) {
        auto at = [&](int b) {
            const int bb = h->is_static[f] ? 0 : b;
            return static_cast<cx<T>*>(h->state[bb]) + static_cast<size_t>(f) * h->field_elems;
        };
        p.cur[f] = at(cur);
        p.base[f] = at(base);
        p.next[f] = at(next);
    }
    if (h->fp32) {
        p.cur_k0 = h->k0[cur];
        p.base_k0 = h->k0[base];
        p.next_k0 = h->k0[next];
    }
    return p;
}

hipEvent_t take_event(shhd_gpu* h) {
    if (!h->ev_pool.empty()) {
        hipEvent_t e = h->ev_pool.back();
        h->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

template <typename F>
shhd_status timed_launch(shhd_gpu* h, int cls, F&& launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (h->timing) {
        a = take_event(h);
        b = take_event(h);
        SHHD_HIP(h, hipEventRecord(a, h->stream));
    }
    launch();
    SHHD_HIP(h, hipGetLastError());
    if (h->timing) {
        SHHD_HIP(h, hipEventRecord(b, h->stream));
        h->ev_used.push_back({cls, {a, b}});
    }
    return SHHD_OK;
}

template <typename T>
shhd_status run_stage(shhd_gpu* h, int next, int cur, int base, const StageParams& sp) {
    const Desc& d = h->desc;
    const StagePtrs<T> P = stage_ptrs<T>(h, next, cur, base);
    cx<T>* tb = static_cast<cx<T>*>(h->tbuf);
    cx<T>* ub = static_cast<cx<T>*>(h->ubuf);
    const cx<T>* twx = static_cast<const cx<T>*>(h->twx);
    const cx<T>* twy = static_cast<const cx<T>*>(h->twy);
    const ColPlan& cp = h->cols;
    const RowPlan& rp = h->rows;
    if (d.n_in > 0) {
        K1Fn<T> k1 = pick_k1<T>(cp.NT, cp.V);
        shhd_status st = timed_launch(h, 0, [&] {
            hipLaunchKernelGGL(k1, dim3(cp.ntiles, d.n_in), dim3(cp.NT), cp.lds, h->stream, d, P, tb, twy, cp.C);
        });
        if (st != SHHD_OK) return st;
    }
    if (d.n_in > 0 || d.n_out > 0) {
        K2Fn<T> k2 = pick_k2<T>(rp.NT, rp.V, rp.maxn);
        shhd_status st = timed_launch(h, 1, [&] {
            hipLaunchKernelGGL(k2, dim3(rp.nblocks), dim3(rp.NT), rp.lds, h->stream, d, tb, ub, twx, rp.R, h->dflag);
        });
        if (st != SHHD_OK) return st;
    }
    K3Fn<T> k3 = pick_k3<T>(cp.NT, cp.V, cp.acc);
    return timed_launch(h, 2, [&] {
        hipLaunchKernelGGL(k3, dim3(cp.ntiles, d.ngroups), dim3(cp.NT), cp.lds3, h->stream, d, P, sp,
                           static_cast<const cx<T>*>(ub), twy, cp.C);
    });
}

template <typename T>
shhd_status set_lds_limits(shhd_gpu* h) {
    const ColPlan& cp = h->cols;
    const RowPlan& rp = h->rows;
    SHHD_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(pick_k1<T>(cp.NT, cp.V)),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(cp.lds)));
    SHHD_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(pick_k3<T>(cp.NT, cp.V, cp.acc)),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(cp.lds3)));
    SHHD_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(pick_k2<T>(rp.NT, rp.V, rp.maxn)),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(rp.lds)));
    return SHHD_OK;
}

template <typename T>
shhd_status upload_twiddles(shhd_gpu* h) {
    std::vector<cx<T>> tx, ty;
    make_twiddles<T>(h->cfg.nx, tx);
    make_twiddles<T>(h->cfg.ny, ty);
    SHHD_HIP(h, hipMalloc(&h->twx, tx.size() * sizeof(cx<T>)));
    SHHD_HIP(h, hipMalloc(&h->twy, ty.size() * sizeof(cx<T>)));
    SHHD_HIP(h, hipMemcpy(h->twx, tx.data(), tx.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
    SHHD_HIP(h, hipMemcpy(h->twy, ty.data(), ty.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
    h->device_bytes += (tx.size() + ty.size()) * sizeof(cx<T>);
    return SHHD_OK;
}

size_t ref_spec(const shhd_config& c) { return static_cast<size_t>(c.nx / 2 + 1) * c.ny; }

bool ref_active(const shhd_config& c, int gx, int gy) {
    if (gx < 0 || gx >= c.nx_active / 2) return false;
    const int ky = (gy <= c.ny / 2) ? gy : gy - c.ny;
    return std::abs(ky) < c.ny_active / 2;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

shhd_status shhd_gpu_create(const shhd_config* config, shhd_gpu** out) {
    if (!config || !out) return set_err(nullptr, SHHD_ERR_INVALID, "null argument");
    *out = nullptr;
    shhd_status st = validate(*config);
    if (st != SHHD_OK) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return set_err(nullptr, SHHD_ERR_HIP, "no HIP device available");

    shhd_gpu* h = new shhd_gpu;
    h->cfg = *config;
    h->fp32 = config->precision == SHHD_PRECISION_FP32;
    h->elem = h->fp32 ? sizeof(cx<float>) : sizeof(cx<double>);
    h->stages = config->integrator == SHHD_INTEGRATOR_SRK3 ? 3 : 1;
    auto fail = [&](shhd_status s) {
        g_create_error = h->err;
        shhd_gpu_destroy(h);
        return s;
    };
    if (hipGetDevice(&h->device) != hipSuccess) return fail(set_err(h, SHHD_ERR_HIP, "hipGetDevice failed"));
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(set_err(h, SHHD_ERR_HIP, "hipStreamCreate failed"));
    st = build_desc(h);
    if (st != SHHD_OK) return fail(st);

    const Desc& d = h->desc;
    h->cols = plan_columns(d.my, d.nay, d.hxa, h->elem);
    h->rows = plan_rows(d.mx, d.my, h->elem);
    int npsi_roles = 0;
    for (int i = 0; i < d.n_in; ++i)
        if (d.in_role[i] >= ROLE_PSI0) npsi_roles = std::max(npsi_roles, d.in_role[i] - ROLE_PSI0 + 1);
    h->rows.maxn = npsi_roles == 0 ? 0 : (npsi_roles == 1 ? 1 : (npsi_roles == 2 ? 2 : (npsi_roles <= 4 ? 4 : 8)));

    h->field_elems = static_cast<size_t>(d.hxa) * d.nay;
    h->plane_elems = static_cast<size_t>(d.my) * d.hxa;
    const size_t nf = static_cast<size_t>(d.nfields);
    for (int b = 0; b < 3; ++b) {
        if (hipMalloc(&h->state[b], nf * h->field_elems * h->elem) != hipSuccess)
            return fail(set_err(h, SHHD_ERR_HIP, "hipMalloc(state) failed"));
        if (hipMemset(h->state[b], 0, nf * h->field_elems * h->elem) != hipSuccess)
            return fail(set_err(h, SHHD_ERR_HIP, "hipMemset(state) failed"));
        h->device_bytes += nf * h->field_elems * h->elem;
        if (h->fp32) {
            if (hipMalloc(&h->k0[b], nf * sizeof(cx<double>)) != hipSuccess)
                return fail(set_err(h, SHHD_ERR_HIP, "hipMalloc(k0) failed"));
            if (hipMemset(h->k0[b], 0, nf * sizeof(cx<double>)) != hipSuccess)
                return fail(set_err(h, SHHD_ERR_HIP, "hipMemset(k0) failed"));
        }
    }
    const size_t tbytes = std::max<size_t>(1, static_cast<size_t>(d.n_in)) * h->plane_elems * h->elem;
    const size_t ubytes = std::max<size_t>(1, static_cast<size_t>(d.n_out)) * h->plane_elems * h->elem;
    if (hipMalloc(&h->tbuf, tbytes) != hipSuccess || hipMalloc(&h->ubuf, ubytes) != hipSuccess)
        return fail(set_err(h, SHHD_ERR_HIP, "hipMalloc(work) failed"));
    h->device_bytes += tbytes + ubytes;
    if (hipMalloc(&h->dflag, sizeof(int)) != hipSuccess) return fail(set_err(h, SHHD_ERR_HIP, "hipMalloc(flag) failed"));
    if (!h->static_fields.empty()) {
        if (hipMalloc(&h->dstatic, h->static_fields.size() * sizeof(int)) != hipSuccess)
            return fail(set_err(h, SHHD_ERR_HIP, "hipMalloc(static) failed"));
        if (hipMemcpy(h->dstatic, h->static_fields.data(), h->static_fields.size() * sizeof(int),
                      hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(h, SHHD_ERR_HIP, "hipMemcpy(static) failed"));
    }
    st = h->fp32 ? upload_twiddles<float>(h) : upload_twiddles<double>(h);
    if (st != SHHD_OK) return fail(st);
    st = h->fp32 ? set_lds_limits<float>(h) : set_lds_limits<double>(h);
    if (st != SHHD_OK) return fail(st);
    if (std::getenv("SHHD_VERBOSE")) {
        std::fprintf(stderr,
                     "[shhd] Mx=%d My=%d Na=%dx%d N=%d fp%d F_in=%d F_out=%d groups=%d | cols C=%d NT=%d V=%d tiles=%d "
                     "lds=%zu/%zu | rows R=%d NT=%d V=%d blocks=%d maxn=%d lds=%zu\n",
                     d.mx, d.my, d.nax, d.nay, d.N, h->fp32 ? 32 : 64, d.n_in, d.n_out, d.ngroups, h->cols.C,
                     h->cols.NT, h->cols.V, h->cols.ntiles, h->cols.lds, h->cols.lds3, h->rows.R, h->rows.NT, h->rows.V,
                     h->rows.nblocks, h->rows.maxn, h->rows.lds);
    }
    *out = h;
    return SHHD_OK;
}

shhd_status shhd_gpu_set_state(shhd_gpu* h, const double* in, size_t n) {
    if (!h || !in) return set_err(h, SHHD_ERR_INVALID, "null argument");
    const shhd_config& c = h->cfg;
    const size_t spec = ref_spec(c);
    const int nf = h->desc.nfields;
    if (n != 2 * spec * static_cast<size_t>(nf)) return set_err(h, SHHD_ERR_INVALID, "state size mismatch");
    const int hx = c.nx / 2 + 1;
    const int nay = c.ny_active;
    std::vector<cx<double>> compact(static_cast<size_t>(nf) * h->field_elems, czero<double>());
    for (int f = 0; f < nf; ++f) {
        const double* src = in + 2 * spec * static_cast<size_t>(f);
        for (int gy = 0; gy < c.ny; ++gy) {
            for (int gx = 0; gx < hx; ++gx) {
                const size_t i = static_cast<size_t>(gy) * hx + gx;
                const double re = src[2 * i], im = src[2 * i + 1];
                if (!ref_active(c, gx, gy)) {
                    if (re != 0.0 || im != 0.0)
                        return set_err(h, SHHD_ERR_INVALID,
                                       "state has nonzero inactive modes (the 3/2-rule mask must be applied)");
                    continue;
                }
                const int ky = (gy <= c.ny / 2) ? gy : gy - c.ny;
                const int cidx = ky >= 0 ? ky : ky + nay;
                compact[static_cast<size_t>(f) * h->field_elems + static_cast<size_t>(gx) * nay + cidx] = mk(re, im);
            }
        }
    }
    SHHD_HIP(h, hipStreamSynchronize(h->stream));
    if (h->fp32) {
        std::vector<cx<float>> cf(compact.size());
        for (size_t i = 0; i < compact.size(); ++i)
            cf[i] = mk(static_cast<float>(compact[i].re), static_cast<float>(compact[i].im));
        SHHD_HIP(h, hipMemcpy(h->state[0], cf.data(), cf.size() * sizeof(cx<float>), hipMemcpyHostToDevice));
        std::vector<cx<double>> k0(static_cast<size_t>(nf));
        for (int f = 0; f < nf; ++f) k0[static_cast<size_t>(f)] = compact[static_cast<size_t>(f) * h->field_elems];
        SHHD_HIP(h, hipMemcpy(h->k0[0], k0.data(), k0.size() * sizeof(cx<double>), hipMemcpyHostToDevice));
    } else {
        SHHD_HIP(h, hipMemcpy(h->state[0], compact.data(), compact.size() * sizeof(cx<double>), hipMemcpyHostToDevice));
    }
    h->cur = 0;
    h->state_set = true;
    // reference density rho_hat(0,0)/(Mx*My) (src/fcalculator_workspace.cc:41-66)
    h->desc.rho0 = compact[0].re / (static_cast<double>(c.nx) * static_cast<double>(c.ny));
    return SHHD_OK;
}

shhd_status shhd_gpu_get_state(shhd_gpu* h, double* out, size_t n) {
    if (!h || !out) return set_err(h, SHHD_ERR_INVALID, "null argument");
    const shhd_config& c = h->cfg;
    const size_t spec = ref_spec(c);
    const int nf = h->desc.nfields;
    if (n != 2 * spec * static_cast<size_t>(nf)) return set_err(h, SHHD_ERR_INVALID, "state size mismatch");
    SHHD_HIP(h, hipStreamSynchronize(h->stream));
    std::vector<cx<double>> compact(static_cast<size_t>(nf) * h->field_elems);
    for (int f = 0; f < nf; ++f) {
        const int b = h->is_static[f] ? 0 : h->cur;
        const size_t off = static_cast<size_t>(f) * h->field_elems;
        if (h->fp32) {
            std::vector<cx<float>> cf(h->field_elems);
            SHHD_HIP(h, hipMemcpy(cf.data(), static_cast<cx<float>*>(h->state[b]) + off, cf.size() * sizeof(cx<float>),
                                  hipMemcpyDeviceToHost));
            for (size_t i = 0; i < cf.size(); ++i) compact[off + i] = mk<double>(cf[i].re, cf[i].im);
            cx<double> k0v;
            SHHD_HIP(h, hipMemcpy(&k0v, h->k0[b] + f, sizeof(cx<double>), hipMemcpyDeviceToHost));
            compact[off] = k0v;
        } else {
            SHHD_HIP(h, hipMemcpy(compact.data() + off, static_cast<cx<double>*>(h->state[b]) + off,
                                  h->field_elems * sizeof(cx<double>), hipMemcpyDeviceToHost));
        }
    }
    std::memset(out, 0, sizeof(double) * n);
    const int hx = c.nx / 2 + 1;
    const int nay = c.ny_active;
    for (int f = 0; f < nf; ++f) {
        double* dst = out + 2 * spec * static_cast<size_t>(f);
        for (int gy = 0; gy < c.ny; ++gy) {
            for (int gx = 0; gx < c.nx_active / 2; ++gx) {
                if (!ref_active(c, gx, gy)) continue;
                const int ky = (gy <= c.ny / 2) ? gy : gy - c.ny;
                const int cidx = ky >= 0 ? ky : ky + nay;
                const cx<double> v = compact[static_cast<size_t>(f) * h->field_elems + static_cast<size_t>(gx) * nay + cidx];
                const size_t i = static_cast<size_t>(gy) * hx + gx;
                dst[2 * i] = v.re;
                dst[2 * i + 1] = v.im;
            }
        }
    }
    return SHHD_OK;
}

shhd_status shhd_gpu_step(shhd_gpu* h, int nsteps, double t0) {
    (void)t0;   // stage times are cache keys only in the reference
    if (!h) return set_err(h, SHHD_ERR_INVALID, "null handle");
    if (!h->state_set) return set_err(h, SHHD_ERR_INVALID, "shhd_gpu_set_state must be called before stepping");
    if (nsteps < 0) return set_err(h, SHHD_ERR_INVALID, "nsteps must be nonnegative");
    if (nsteps == 0) return SHHD_OK;
    if (h->rho0_needed && h->desc.rho0 == 0.0) return set_err(h, SHHD_ERR_RUNTIME, "reference density must be nonzero.");
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, h->stream, h->dflag);
    if (!h->static_fields.empty()) {
        const int ns = static_cast<int>(h->static_fields.size());
        if (h->fp32)
            hipLaunchKernelGGL(k_enforce_static<float>, dim3(ns), dim3(256), 0, h->stream,
                               static_cast<cx<float>*>(h->state[0]), h->desc.nay, h->field_elems, h->dstatic, ns, h->k0[0]);
        else
            hipLaunchKernelGGL(k_enforce_static<double>, dim3(ns), dim3(256), 0, h->stream,
                               static_cast<cx<double>*>(h->state[0]), h->desc.nay, h->field_elems, h->dstatic, ns,
                               static_cast<cx<double>*>(nullptr));
        SHHD_HIP(h, hipGetLastError());
    }
    static constexpr double S2 = 1.41421356237309504880;
    static constexpr double S3 = 1.73205080756887729352;
    static constexpr double B1 = (2.0 * S2 + S3) / 5.0;
    static constexpr double B2 = (-4.0 * S2 + 3.0 * S3) / 5.0;
    static constexpr double B3 = (S2 - 2.0 * S3) / 10.0;
    for (int s = 0; s < nsteps; ++s) {
        StageParams sp{};
        sp.step_lo = static_cast<uint32_t>(h->step_index);
        sp.step_hi = static_cast<uint32_t>(h->step_index >> 32);
        const int x = h->cur, y = (h->cur + 1) % 3, z = (h->cur + 2) % 3;
        shhd_status st = SHHD_OK;
        auto stage = [&](int next, int cur, int base, double beta, double wb, double wc, int use_b) {
            sp.beta = beta;
            sp.wb = wb;
            sp.wc = wc;
            sp.use_b = use_b;
            return h->fp32 ? run_stage<float>(h, next, cur, base, sp) : run_stage<double>(h, next, cur, base, sp);
        };
        if (h->stages == 3) {
            // time_integrator_srk3_*.h: (w_base, w_current) = (0,1), (3/4,1/4), (1/3,2/3)
            if ((st = stage(y, x, x, B1, 0.0, 1.0, 1)) != SHHD_OK) return st;
            if ((st = stage(z, y, x, B2, 3.0 / 4.0, 1.0 / 4.0, 1)) != SHHD_OK) return st;
            if ((st = stage(y, z, x, B3, 1.0 / 3.0, 2.0 / 3.0, 1)) != SHHD_OK) return st;
            h->cur = y;
        } else {
            if ((st = stage(y, x, x, 0.0, 0.0, 1.0, 0)) != SHHD_OK) return st;
            h->cur = y;
        }
        h->step_index++;
    }
    int flag = 0;
    SHHD_HIP(h, hipMemcpyAsync(&flag, h->dflag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    SHHD_HIP(h, hipStreamSynchronize(h->stream));
    if (h->timing) {
        for (auto& u : h->ev_used) {
            float ms = 0.f;
            SHHD_HIP(h, hipEventElapsedTime(&ms, u.second.first, u.second.second));
            h->time_ms[u.first] += ms;
            h->launches[u.first] += 1;
            h->ev_pool.push_back(u.second.first);
            h->ev_pool.push_back(u.second.second);
        }
        h->ev_used.clear();
    }
    if (flag) return set_err(h, SHHD_ERR_ZERO_DENSITY, "compressible velocity encountered zero density.");
    return SHHD_OK;
}

shhd_status shhd_gpu_set_flux_request(shhd_gpu* h, int density, int order_parameter, int momentum) {
    if (!h) return set_err(h, SHHD_ERR_INVALID, "null handle");
    if (density || order_parameter || momentum)
        return set_err(h, SHHD_ERR_UNSUPPORTED, "flux recording has no device path");
    return SHHD_OK;
}

shhd_status shhd_gpu_get_flux(shhd_gpu* h, double* out, size_t n) {
    (void)out;
    (void)n;
    return set_err(h, SHHD_ERR_UNSUPPORTED, "flux recording has no device path");
}

void shhd_gpu_destroy(shhd_gpu* h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (int b = 0; b < 3; ++b) {
        if (h->state[b]) (void)hipFree(h->state[b]);
        if (h->k0[b]) (void)hipFree(h->k0[b]);
    }
    if (h->tbuf) (void)hipFree(h->tbuf);
    if (h->ubuf) (void)hipFree(h->ubuf);
    if (h->twx) (void)hipFree(h->twx);
    if (h->twy) (void)hipFree(h->twy);
    if (h->dflag) (void)hipFree(h->dflag);
    if (h->dstatic) (void)hipFree(h->dstatic);
    for (auto e : h->ev_pool) (void)hipEventDestroy(e);
    for (auto& u : h->ev_used) {
        (void)hipEventDestroy(u.second.first);
        (void)hipEventDestroy(u.second.second);
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* shhd_gpu_last_error(const shhd_gpu* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

shhd_status shhd_gpu_get_info(const shhd_gpu* h, shhd_info* info) {
    if (!h || !info) return SHHD_ERR_INVALID;
    info->num_fields = h->desc.nfields;
    info->num_inverse_fields = h->n_in;
    info->num_forward_fields = h->n_out;
    info->num_updated_fields = h->n_updated;
    info->stages_per_step = h->stages;
    info->state_doubles = 2 * ref_spec(h->cfg) * static_cast<size_t>(h->desc.nfields);
    info->device_bytes = h->device_bytes;
    return SHHD_OK;
}

shhd_status shhd_gpu_set_kernel_timing(shhd_gpu* h, int enable) {
    if (!h) return SHHD_ERR_INVALID;
    h->timing = enable != 0;
    for (int i = 0; i < kClasses; ++i) {
        h->time_ms[i] = 0.0;
        h->launches[i] = 0;
    }
    return SHHD_OK;
}

shhd_status shhd_gpu_get_kernel_timing(shhd_gpu* h, double* total_ms, long long* launches, int nclasses) {
    if (!h || !total_ms || !launches) return SHHD_ERR_INVALID;
    for (int i = 0; i < nclasses && i < kClasses; ++i) {
        total_ms[i] = h->time_ms[i];
        launches[i] = h->launches[i];
    }
    return SHHD_OK;
}

shhd_status shhd_gpu_get_kernel_bytes(const shhd_gpu* h, double* bytes, int nclasses) {
    if (!h || !bytes) return SHHD_ERR_INVALID;
    const Desc& d = h->desc;
    const double fb = static_cast<double>(h->field_elems) * static_cast<double>(h->elem);
    const double pb = static_cast<double>(h->plane_elems) * static_cast<double>(h->elem);
    double k3 = d.n_out * pb;
    for (int g = 0; g < d.ngroups; ++g) {
        const Group& G = d.groups[g];
        if (G.kind == GROUP_PSI) {
            k3 += 3.0 * fb;   // cur, base, next of psi_a
            k3 += static_cast<double>(d.n_igrad > 0 ? 2 : 0) * fb;
        } else if (d.mode == SHHD_MODE_COMPRESSIBLE) {
            k3 += 9.0 * fb;   // rho, jx, jy: cur, base, next
        } else {
            k3 += 6.0 * fb + (d.pressure_coefficient != 0.0 ? fb : 0.0);
        }
    }
    const double vals[kClasses] = {d.n_in * (fb + pb), (d.n_in + d.n_out) * pb, k3};
    for (int i = 0; i < nclasses && i < kClasses; ++i) bytes[i] = vals[i];
    return SHHD_OK;
}

}  // extern "C"
