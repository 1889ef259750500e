// shhd_gpu.hip — host side of the MI355X right-hand-side and time-step engine:
// descriptor lowering, launch planning, device tables, stepping (eager or
// replayed from a captured hipGraph) and the C-ABI of include/shhd_gpu.h.
//
// Every simulation is an Instance.  Its state lives in HBM in a compact
// spectral layout holding active modes only:
//     field f, column gx < Nax/2, slot c < Nay  ->  index f*(Nax/2*Nay) + gx*Nay + c
// with ky = c for c < Nay/2 and ky = c - Nay otherwise (slot Nay/2 is the
// always-zero ky = -Nay/2 mode).  Per stage three kernels run:
//     K1 (inverse y, F_in fields) -> K2 (rows: x transforms + products)
//     -> K3 (forward y + spectral RHS + noise + stage update).
// An Engine holds any number of instances; instances with the same launch
// shape form a class and each kernel of a stage is ONE launch per class, the
// workgroups being mapped to (instance, local block) through device tables.
// A single-simulation handle (shhd_gpu) is an engine with one instance; the
// batch API (shhd_batch) exposes the engine for ensembles and serving.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/shhd_gpu.h"
#include "instances.h"
#include "measure.h"
#include "ic.h"

using namespace shhd;

// ---------------------------------------------------------------------------
// Scratch of the entry point that allocates nothing
// (shhd_batch_apply_initial_condition_noalloc).  The velocity normalization of
// the random_vorticity and spectrum/k4 initial conditions (k_ic_norm) writes
// one FP64 partial sum per workgroup, at most kIcNormBlocks of them whatever
// the grid; the allocating entry point takes them from a per-call hipMalloc,
// the other from this array.  It holds 2048 doubles (16 KiB) on every device
// the library runs on, until the process ends, and keeps the last call's
// partials: a call writes every partial it reads before reading it, so
// nothing is cleared.  Two such calls that overlap on one device race on it.
// In host code `scratch` is a host-side stand-in, not a device address: its
// address on each device comes from hipGetSymbolAddress (scratch_address
// below), and only that address reaches kernels and copies.
constexpr int kIcNormBlocks = 2048;
static __device__ __align__(256) double scratch[kIcNormBlocks];

namespace {

std::string g_create_error;

constexpr int kClasses = 3;
constexpr size_t kLdsCap = 96 * 1024;
constexpr size_t kLdsCap3 = 128 * 1024;
// Two transform buffers only while the workgroups a CU can hold by their waves
// still fit its 160 KiB of LDS: two of up to 512 lanes, one of 1024 lanes.
bool two_buffers_fit(size_t one_buffer_bytes, int lanes) {
    return 2 * one_buffer_bytes <= (lanes >= 1024 ? 160 * 1024 : 80 * 1024);
}
constexpr int kMaxSlots = 4096;

// device address of `scratch`, looked up once per device with that device current
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxScratchDevices] = {};

struct ColPlan {
    int C = 1, NT = 256, V = 6, ntiles = 1;
    int acc = 6;       // K3 accumulator slots per lane: 4 when C*Nay <= 4*NT, else V
    int fz = 0;        // column length of the fused build in use, 0 for the staged one
    int nbuf = 1;      // LDS buffers of the fused build
    size_t lds = 0;    // K1: padded C*My values
    size_t lds3 = 0;   // K3: max(padded C*My, 2*C*Nay) values (FFT column, then both accumulators)
    int ws = 1;        // K3: wave groups of 64 lanes splitting the products of a tile (fused builds)
};
struct RowPlan {
    int R = 1, NT = 256, V = 6, nblocks = 1, maxn = 0;
    int fl = 0;        // row length of a fixed-length row-kernel build, 0 for the any-length one
    int nbuf = 0;      // fixed-length build: 0 staged, 1 or 2 fused with that many LDS buffers
    int layout = 0;    // fused build: fixed order of the inverse-transformed fields (0: any)
    size_t lds = 0;
    int ws = 1;        // wave groups of 64 lanes splitting the transforms of a tile (staged builds)
};

// device transform lengths: 2^a or 3 * 2^a (every Stockham sub-length ns is
// then a power of two, which the pass indexing relies on)
bool is_device_length(int n) {
    if (n <= 0) return false;
    if (n % 3 == 0) n /= 3;
    while (n % 2 == 0) n /= 2;
    return n == 1;
}

// Stockham plan of length L: radix 8 passes first (max_radix 8) or radix 4
// (max_radix 4), then 2, then 3.
FftPlanDev make_fft_plan(int L, int max_radix = 8) {
    FftPlanDev p{};
    std::vector<int> rad;
    int m = L;
    while (max_radix >= 8 && m % 8 == 0 && m > 0) { rad.push_back(8); m /= 8; }
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

int lanes_for(int n, int v, int nt_min) {
    int nt = (n <= 64 * v) ? 64 : (n <= 256 * v ? 256 : (n <= 512 * v ? 512 : 1024));
    return std::max(nt, nt_min);
}

// Column tile of C kx columns per workgroup: `c_tile` when the tuner chose
// one, else the tile whose row segments are `seg_bytes` wide; a batch plan
// (nt_min = 256) widens C so small grids share the 256-lane shape of the
// larger ones.  The inverse and forward column passes are planned separately.
ColPlan plan_columns(int my, int nay, int hxa, size_t elem, int nt_min, int seg_bytes, int c_tile) {
    ColPlan p;
    p.V = (my % 3 == 0) ? 6 : 8;
    int target = c_tile > 0 ? c_tile : std::max(1, static_cast<int>(seg_bytes / static_cast<int>(elem)));
    if (nt_min > 0) target = std::max(target, (nt_min * p.V) / my);
    int C = std::min(std::max(target, 1), pow2ceil(hxa));
    while (C & (C - 1)) C &= C - 1;   // a power of two (the column kernels split indices by shifts)
    auto lds3 = [&](int c) { return static_cast<size_t>(std::max(lp_size(c * my), 2 * c * nay)) * elem; };
    while (C > 1 && (lds3(C) > kLdsCap3 || C * my > 1024 * p.V)) C >>= 1;
    p.C = C;
    p.NT = lanes_for(C * my, p.V, nt_min);
    p.ntiles = (hxa + C - 1) / C;
    p.acc = (C * nay <= 4 * p.NT) ? 4 : p.V;
    p.lds = static_cast<size_t>(lp_size(C * my)) * elem;
    p.lds3 = lds3(C);
    return p;
}

int row_fixed_length(int mx, int NT, int V);

RowPlan plan_rows(int mx, int my, size_t elem, int nt_min) {
    RowPlan p;
    p.V = (mx % 3 == 0) ? 6 : 8;
    p.NT = lanes_for(mx, p.V, nt_min);
    p.R = std::max(1, (p.NT * p.V) / mx);
    while (p.R > 1 && static_cast<size_t>(lp_size(p.R * mx)) * elem > kLdsCap) --p.R;
    p.R = std::min(p.R, my);
    p.nblocks = (my + p.R - 1) / p.R;
    p.fl = row_fixed_length(mx, p.NT, p.V);
    if (p.fl && p.R != (p.NT * p.V) / p.fl) p.fl = 0;   // the build assumes the natural rows per workgroup
    p.lds = static_cast<size_t>(lp_size(p.R * mx)) * elem;
    return p;
}

// ---------------------------------------------------------------------------
// kernel dispatch over the instantiated (NT, V, ...) shapes
// ---------------------------------------------------------------------------
template <typename T>
using K1Fn = void (*)(const InstDev<T>*, const StagePtrs<T>*, BlockMap);
template <typename T>
using K2Fn = void (*)(const InstDev<T>*, BlockMap);
template <typename T>
using K3Fn = void (*)(const InstDev<T>*, const StagePtrs<T>*, StageParams, BlockMap);

// forward-column builds with wave groups exist for (NT, V, L)
bool col_ws_built(int NT, int V, int L) {
#define SHHD_CWS(nt, v, len) if (NT == nt && V == v && L == len) return true;
    SHHD_FOR_COL_FZ_WS(SHHD_CWS)
#undef SHHD_CWS
    return false;
}

// row builds with wave groups exist for (NT, V)
bool row_ws_built(int NT, int V) {
#define SHHD_RWS(nt, v) if (NT == nt && V == v) return true;
    SHHD_FOR_ROW_WS(SHHD_RWS)
#undef SHHD_RWS
    return false;
}

// fused column builds exist for (NT, V, L)
bool col_fused_built(int NT, int V, int L) {
#define SHHD_CFZ(nt, v, len) if (NT == nt && V == v && L == len) return true;
    SHHD_FOR_COL_FZ(SHHD_CFZ)
#undef SHHD_CFZ
    return false;
}

template <typename T>
K1Fn<T> pick_k1(int NT, int V, int fz, int nbuf) {
    if (fz > 0) {
#define SHHD_K1FZ(nt, v, len)                                                                       \
    if (NT == nt && V == v && fz == len) return nbuf == 2 ? &k_inverse_columns<T, nt, v, len, 2> \
                                                          : &k_inverse_columns<T, nt, v, len, 1>;
        SHHD_FOR_COL_FZ(SHHD_K1FZ)
#undef SHHD_K1FZ
        return nullptr;
    }
#define SHHD_K1(nt, v) if (NT == nt && V == v) return &k_inverse_columns<T, nt, v, 0, 1>;
    SHHD_FOR_NT_V(SHHD_K1)
#undef SHHD_K1
    return nullptr;
}

template <typename T>
K3Fn<T> pick_k3(int NT, int V, int acc, int fz, int nbuf, int ws) {
    if (fz > 0) {
        if (acc != 2 * V / 3) return nullptr;
        if (ws == 4) {
#define SHHD_K3WS(nt, v, len)                                                                                     \
    if (NT == nt && V == v && fz == len) return nbuf == 2 ? &k_forward_columns<T, nt, v, (2 * (v) / 3), len, 2, 4> \
                                                          : &k_forward_columns<T, nt, v, (2 * (v) / 3), len, 1, 4>;
            SHHD_FOR_COL_FZ_WS(SHHD_K3WS)
#undef SHHD_K3WS
            return nullptr;
        }
        if (ws != 1) return nullptr;
#define SHHD_K3FZ(nt, v, len)                                                                                     \
    if (NT == nt && V == v && fz == len) return nbuf == 2 ? &k_forward_columns<T, nt, v, (2 * (v) / 3), len, 2, 1> \
                                                          : &k_forward_columns<T, nt, v, (2 * (v) / 3), len, 1, 1>;
        SHHD_FOR_COL_FZ(SHHD_K3FZ)
#undef SHHD_K3FZ
        return nullptr;
    }
    if (ws != 1) return nullptr;
#define SHHD_K3(nt, v) \
    if (NT == nt && V == v && acc == 4) return &k_forward_columns<T, nt, v, 4, 0, 1, 1>; \
    if (NT == nt && V == v && acc == v) return &k_forward_columns<T, nt, v, v, 0, 1, 1>;
    SHHD_FOR_NT_V(SHHD_K3)
#undef SHHD_K3
    return nullptr;
}

// row length with its own row-kernel build for this lane shape, else 0
int row_fixed_length(int mx, int NT, int V) {
#define SHHD_RFL(nt, v, len) if (mx == len && NT == nt && V == v) return len;
    SHHD_FOR_ROW_FL(SHHD_RFL)
#undef SHHD_RFL
    return 0;
}

// a fused row build with a fixed field layout exists
bool row_layout_built(int NT, int V, int maxn, int fl, int nbuf, int layout) {
#define SHHD_RLN(nt, v, len, nb, lay, mn) \
    if (NT == nt && V == v && fl == len && nbuf == nb && layout == lay && maxn == mn) return true;
#define SHHD_RL(nt, v, len) SHHD_FOR_LAYOUT(SHHD_RLN, nt, v, len, 1) SHHD_FOR_LAYOUT(SHHD_RLN, nt, v, len, 2)
    SHHD_FOR_ROW_FL(SHHD_RL)
#undef SHHD_RL
#undef SHHD_RLN
    return false;
}

template <typename T>
K2Fn<T> pick_k2(int NT, int V, int maxn, int fl, int nbuf, int layout, int ws) {
    const int n = maxn <= 0 ? 0 : (maxn == 1 ? 1 : (maxn == 2 ? 2 : (maxn <= 4 ? 4 : 8)));
    if (ws == 4) {
        if (fl != 0 || nbuf != 0 || layout != 0) return nullptr;
#define SHHD_K2WN(nt, v, mn) if (NT == nt && V == v && n == mn) return &k_rows<T, nt, v, mn, 0, 0, 0, 4>;
#define SHHD_K2W(nt, v) SHHD_FOR_MAXN(SHHD_K2WN, nt, v)
        SHHD_FOR_ROW_WS(SHHD_K2W)
#undef SHHD_K2W
#undef SHHD_K2WN
        return nullptr;
    }
    if (ws != 1) return nullptr;
    if (fl > 0 && nbuf > 0 && layout > 0) {
#define SHHD_K2LN(nt, v, len, nb, lay, mn) \
    if (NT == nt && V == v && fl == len && nbuf == nb && layout == lay && n == mn) return &k_rows<T, nt, v, mn, len, nb, lay, 1>;
#define SHHD_K2L(nt, v, len) SHHD_FOR_LAYOUT(SHHD_K2LN, nt, v, len, 1) SHHD_FOR_LAYOUT(SHHD_K2LN, nt, v, len, 2)
        SHHD_FOR_ROW_FL(SHHD_K2L)
#undef SHHD_K2L
#undef SHHD_K2LN
    }
    if (fl > 0) {
#define SHHD_K2FN(nt, v, len, mn)                                                  \
    if (NT == nt && V == v && n == mn && fl == len)                                \
        return nbuf == 2 ? &k_rows<T, nt, v, mn, len, 2, 0, 1>                     \
                         : (nbuf == 1 ? &k_rows<T, nt, v, mn, len, 1, 0, 1> : &k_rows<T, nt, v, mn, len, 0, 0, 1>);
#define SHHD_K2F(nt, v, len) SHHD_FOR_MAXN(SHHD_K2FN, nt, v, len)
        SHHD_FOR_ROW_FL(SHHD_K2F)
#undef SHHD_K2F
#undef SHHD_K2FN
    }
#define SHHD_K2N(nt, v, mn) if (NT == nt && V == v && n == mn) return &k_rows<T, nt, v, mn, 0, 0, 0, 1>;
#define SHHD_K2(nt, v) SHHD_K2N(nt, v, 0) SHHD_K2N(nt, v, 1) SHHD_K2N(nt, v, 2) SHHD_K2N(nt, v, 4) SHHD_K2N(nt, v, 8)
    SHHD_FOR_NT_V(SHHD_K2)
#undef SHHD_K2
#undef SHHD_K2N
    return nullptr;
}

template <typename T>
using KPFn = void (*)(const InstDev<T>*, T*, int);

template <typename T>
KPFn<T> pick_kp(int NT, int V) {
#define SHHD_KP(nt, v) if (NT == nt && V == v) return &k_rows_physical<T, nt, v>;
    SHHD_FOR_NT_V(SHHD_KP)
#undef SHHD_KP
    return nullptr;
}

__global__ void k_clear_flags(int* flags, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = 0;
}

// step counters of the instances that completed the step: those in a launch
// class (active) whose zero-density flag is clear
__global__ void k_advance(uint64_t* steps, const int* flags, const int* active, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && active[i] != 0 && flags[i] == 0) steps[i] += 1;
}

// kx = 0 column of a field that no stage rewrites (rho when incompressible,
// rho and j when quiescent): Hermitian projection + real self-conjugate mode.
// Idempotent, so it runs at the start of every step.
template <typename T>
struct StaticEntry {
    cx<T>* col;
    cx<double>* k0;   // FP64 copy of the k = 0 mode (FP32 form) or null
    int nay;
};

template <typename T>
__global__ void k_enforce_static(const StaticEntry<T>* __restrict__ entries) {
    const StaticEntry<T> e = entries[blockIdx.x];
    const int half = e.nay / 2;
    for (int m = threadIdx.x; m < half; m += blockDim.x) {
        if (m == 0) {
            e.col[0].im = T(0);
            if (e.k0) e.k0->im = 0.0;
        } else {
            const cx<T> a = e.col[m];
            const cx<T> b = e.col[e.nay - m];
            const cx<T> p = T(0.5) * (a + conj(b));
            e.col[m] = p;
            e.col[e.nay - m] = conj(p);
        }
    }
}

shhd_status set_err(std::string* target, shhd_status st, const std::string& msg) {
    if (target) *target = msg;
    else g_create_error = msg;
    return st;
}
shhd_status set_err(std::nullptr_t, shhd_status st, const std::string& msg) {
    g_create_error = msg;
    return st;
}

#define SHHD_HIP(errp, call)                                                                      \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return set_err(errp, SHHD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// The device address of `scratch` on `device`, which must be the current
// device (hipGetSymbolAddress answers for the current one).
shhd_status scratch_address(int device, double** out, std::string* err) {
    int cur = -1;
    SHHD_HIP(err, hipGetDevice(&cur));
    if (cur != device) return set_err(err, SHHD_ERR_HIP, "the instance's device is not the current device");
    if (device < 0 || device >= kMaxScratchDevices)
        return set_err(err, SHHD_ERR_HIP, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        SHHD_HIP(err, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_err(err, SHHD_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[device];
    return SHHD_OK;
}

template <typename T>
void make_twiddles(int L, std::vector<cx<T>>& out) {
    out.resize(static_cast<size_t>(fz_tw_size(L)));
    const double pi = 3.14159265358979323846;
    for (int j = 0; j < L; ++j) {
        const double a = 2.0 * pi * static_cast<double>(j) / static_cast<double>(L);
        out[static_cast<size_t>(j)] = mk(static_cast<T>(std::cos(a)), static_cast<T>(-std::sin(a)));
    }
    // the radix-8 twiddle planes of the fused transforms (layout in fft_fused.h)
    for (const bool r3f : {false, true}) {
        for (int P = 1; P < fz_npass(L, r3f); ++P) {
            if (fz_radix(L, r3f, P) != 8) continue;
            const int ns = fz_ns(L, r3f, P);
            const int tstep = L / (8 * ns);
            const size_t at = static_cast<size_t>(fz_tw8_at(L, r3f, P));
            for (int k = 0; k < ns; ++k)
                for (int q = 1; q < 8; ++q) {
                    const size_t e = sizeof(T) == 4 ? 2 * static_cast<size_t>(((q - 1) / 2) * ns + k) + (q - 1) % 2
                                                    : static_cast<size_t>((q - 1) * ns + k);
                    out[at + e] = out[static_cast<size_t>(q * k * tstep)];
                }
        }
    }
}

size_t ref_spec(const shhd_config& c) { return static_cast<size_t>(c.nx / 2 + 1) * c.ny; }

bool ref_active(const shhd_config& c, int gx, int gy) {
    if (gx < 0 || gx >= c.nx_active / 2) return false;
    const int ky = (gy <= c.ny / 2) ? gy : gy - c.ny;
    return std::abs(ky) < c.ny_active / 2;
}

// ---------------------------------------------------------------------------
// one simulation
// ---------------------------------------------------------------------------
struct Instance {
    shhd_config cfg{};
    Desc desc{};
    bool fp32 = false;
    size_t elem = 16;
    int stages = 3;
    ColPlan cols;     // inverse-column pass (K1, physical export)
    ColPlan cols3;    // forward-column pass (K3)
    RowPlan rows;
    int n_in = 0, n_out = 0, n_updated = 0;
    size_t field_elems = 0;   // hxa * nay
    size_t plane_elems = 0;   // my * hxa
    void* state[3] = {nullptr, nullptr, nullptr};
    cx<double>* k0[3] = {nullptr, nullptr, nullptr};
    void* tbuf = nullptr;
    void* ubuf = nullptr;
    void* twx = nullptr;
    void* twy = nullptr;
    // state, k0, tbuf, ubuf and the twiddle tables above are carved out
    // of this one allocation (one hipMalloc and one hipFree per instance)
    void* slab = nullptr;
    std::vector<int> static_fields;
    bool is_static[kMaxFields] = {};
    bool rho0_needed = false;
    int cur = 0;
    uint64_t step_index = 0;
    // the density reached zero in step `step_index` (0-based): the state is
    // that step's start and the instance takes no further step
    bool failed = false;
    // held by the caller (shhd_batch_hold): keeps its state, takes no step
    bool held = false;
    size_t bytes = 0;
    bool batch_shape = false;
    std::string tune_key;
    // flux recording (FluxBuffer, src/buffer_flux.h): slots of field_elems each
    void* flux = nullptr;
    int flux_req = 0;           // FluxBits
    bool flux_clear = false;    // zero every slot before the next step
    int flux_slots() const { return 2 + 2 * desc.N + 3; }
    // on-device accumulators (measure.h)
    std::vector<Probe> probes;
    Probe* d_probes = nullptr;
    cx<double>* d_ring = nullptr;
    int* d_rows = nullptr;
    int capacity = 0;
    int probe_rows = 0;         // rows recorded since the last read (host count)
    size_t probe_bytes = 0;
    std::vector<int> sk_fields;
    int sk_nevery = 0;
    uint64_t sk_base = 0;
    double* d_sk = nullptr;
    unsigned long long* d_sk_count = nullptr;
    size_t sk_bytes = 0;

    void free_probes() {
        if (d_probes) (void)hipFree(d_probes);
        if (d_ring) (void)hipFree(d_ring);
        if (d_rows) (void)hipFree(d_rows);
        d_probes = nullptr;
        d_ring = nullptr;
        d_rows = nullptr;
        probes.clear();
        capacity = probe_rows = 0;
    }
    void free_sk() {
        if (d_sk) (void)hipFree(d_sk);
        if (d_sk_count) (void)hipFree(d_sk_count);
        d_sk = nullptr;
        d_sk_count = nullptr;
        sk_fields.clear();
        sk_nevery = 0;
    }

    ~Instance() {
        free_probes();
        free_sk();
        if (flux) (void)hipFree(flux);
        if (slab) (void)hipFree(slab);
    }
};

// Coefficients of one product's spectral terms (ContribCoef in desc.h):
//   mu_phys:   F = -i M k mu_hat,            rhs -= M k^2 mu_hat  (src/fcalculator.cc:244-256)
//   psi v:     F = (psi v)_hat,               rhs -= i k . F       (258-272)
//   gradient:  rhs += source (compressible force/gradient, 776-840); no flux
//   v_x, v_y:  Pi = -i (2 eta k v - (eta - zeta)(k . v) 1), rhs -= i k . Pi  (343-398)
//   j v:       Pi = (j v)_hat,                 rhs -= i k . Pi      (472-491)
ContribCoef contrib_coef(const Desc& d, OutItem it) {
    ContribCoef c{};
    const double eta = d.eta, zeta = d.zeta;
    switch (it.kind) {
        case OUT_MU: {
            const double M = d.mobility[it.alpha];
            c.a[0][1] = -M;
            c.a[0][2] = -M;
            c.f[0][1] = M;        // F_x = -i M kx mu
            c.f[1][2] = M;        // F_y = -i M ky mu
            break;
        }
        case OUT_PSIVX: c.b[0][0] = 1.0; c.f[0][0] = 1.0; break;
        case OUT_PSIVY: c.b[0][1] = 1.0; c.f[1][0] = 1.0; break;
        case OUT_GRADF: c.a[0][0] = 1.0; break;
        case OUT_VX:
            // Pi_xx = -i (eta + zeta) kx v, Pi_xy = -i eta ky v, Pi_yy = i (eta - zeta) kx v
            c.a[0][1] = -(eta + zeta);
            c.a[0][2] = -eta;
            c.a[1][3] = -zeta;
            c.f[0][1] = eta + zeta;
            c.f[1][2] = eta;
            c.f[2][1] = -(eta - zeta);
            break;
        case OUT_VY:
            // Pi_xx = i (eta - zeta) ky v, Pi_xy = -i eta kx v, Pi_yy = -i (eta + zeta) ky v
            c.a[0][3] = -zeta;
            c.a[1][1] = -eta;
            c.a[1][2] = -(eta + zeta);
            c.f[0][2] = -(eta - zeta);
            c.f[1][1] = eta;
            c.f[2][2] = eta + zeta;
            break;
        case OUT_PXX: c.b[0][0] = 1.0; c.f[0][0] = 1.0; break;
        case OUT_PXY: c.b[0][1] = 1.0; c.b[1][0] = 1.0; c.f[1][0] = 1.0; break;
        default: c.b[1][1] = 1.0; c.f[2][0] = 1.0; break;   // OUT_PYY
    }
    c.nz = 0;
    for (int r = 0; r < 2; ++r)
        for (int t = 0; t < 4; ++t)
            if (c.a[r][t] != 0.0) c.nz |= 1u << (4 * r + t);
    for (int r = 0; r < 2; ++r)
        for (int t = 0; t < 2; ++t)
            if (c.b[r][t] != 0.0) c.nz |= 1u << (8 + 2 * r + t);
    for (int t = 0; t < 3; ++t)
        for (int u = 0; u < 3; ++u)
            if (c.f[t][u] != 0.0) c.nz |= 1u << (12 + 3 * t + u);
    for (int r = 0; r < 2; ++r) {
        for (int t = 0; t < 4; ++t) c.af[r][t] = static_cast<float>(c.a[r][t]);
        for (int t = 0; t < 2; ++t) c.bf[r][t] = static_cast<float>(c.b[r][t]);
    }
    for (int t = 0; t < 3; ++t)
        for (int u = 0; u < 3; ++u) c.ff[t][u] = static_cast<float>(c.f[t][u]);
    return c;
}

// Lower shhd_config into the kernel descriptor (see desc.h for the meaning of
// each entry).  Mirrors the reference's request logic:
// PhysicalRHSFieldRequests (src/fcalculator_field_requests_and_layout.h:30-127),
// the early returns of psi_det / j_det (src/fcalculator.cc:86-150) and the
// noise enablement (src/fluid_solver.h:79-94).
shhd_status build_desc(Instance* h) {
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
    auto log2_or = [](int n) {
        int l = 0;
        while ((1 << l) < n) ++l;
        return (1 << l) == n ? l : -1;
    };
    d.lnay = log2_or(d.nay);
    d.lhxa = log2_or(d.hxa);
    d.dkx = 2.0 * 3.14159265358979323846 / c.lx;
    d.dky = 2.0 * 3.14159265358979323846 / c.ly;
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
        for (int i = 0; i < g.ncontrib; ++i) {
            g.contrib[i] = psi_contrib[a][static_cast<size_t>(i)];
            g.coef[i] = contrib_coef(d, d.out[g.contrib[i]]);
        }
    }
    if (!quiescent) {
        Group& g = d.groups[d.ngroups++];
        g.kind = GROUP_MOM;
        g.alpha = 0;
        g.ncontrib = static_cast<int>(mom_contrib.size());
        for (int i = 0; i < g.ncontrib; ++i) {
            g.contrib[i] = mom_contrib[static_cast<size_t>(i)];
            g.coef[i] = contrib_coef(d, d.out[g.contrib[i]]);
        }
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


// ---------------------------------------------------------------------------
// launch-shape tuning: the column tiles of the two column passes and the lanes
// of the row pass, measured by shhd_autotune for a shape of problem and kept
// for the process (every later instance of that shape uses them)
// ---------------------------------------------------------------------------
struct Tuning {
    int c1 = 0, c3 = 0, rnt = 0;   // 0: the untuned default
    // transforms: 0 fixed-length build where one exists (default), 1 runtime
    // plan with radix-8 passes first, 2 runtime plan with radix-4 passes,
    // 3 and 4 fused fixed-length build with one and two LDS buffers (else 0)
    int col_transform = 0, row_transform = 0;
    // U planes in blocks of 2^ulg kx columns (0: column-major); a block
    // that does not divide Nax/2 falls back to 0
    int ulg = 0;
    // T planes: -1 row-major, else blocks of 2^tlg kx columns (measured: the
    // blocks pay where the inverse column pass writes one narrow tile per
    // workgroup into long rows, e.g. FP32 at 4096^2)
    int tlg = -1;
    // forward column pass and row pass: 0 wave groups where apply_plans'
    // rule picks them, 1 never, 4 wherever a build exists
    int k3ws = 0;
    int k2ws = 0;
};
// launch shapes of a shape that was never tuned: default tiles and lanes,
// the fused transforms with two LDS buffers where they fit (else one), U
// blocks of four columns (batch instances too: measured faster for the
// mixed-size ensemble and the served stream than the staged builds, although
// a batch then launches once per transform length)
Tuning default_tuning() {
    Tuning t;
    t.col_transform = 4;
    t.row_transform = 4;
    t.ulg = 2;
    return t;
}
constexpr int kTransforms = 5;
std::mutex g_tune_mutex;
std::map<std::string, Tuning> g_tuned;

// what the launch shapes depend on: precision, grids, transform counts,
// groups, stages, mode, order parameters held in the row pass, noise, batch plan
std::string tune_key_of(const Instance& h) {
    const Desc& d = h.desc;
    bool noise = d.j_noise_on != 0;
    for (int a = 0; a < d.N; ++a) noise = noise || d.psi_noise_on[a] != 0;
    char buf[256];
    std::snprintf(buf, sizeof(buf), "fp%d %dx%d a%dx%d in%d out%d g%d s%d m%d n%d z%d b%d", h.fp32 ? 32 : 64, d.mx,
                  d.my, d.nax, d.nay, d.n_in, d.n_out, d.ngroups, h.stages, d.mode, d.N, noise ? 1 : 0,
                  h.batch_shape ? 1 : 0);
    return buf;
}

bool lookup_tuning(const std::string& key, Tuning& t) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tuned.find(key);
    if (it == g_tuned.end()) return false;
    t = it->second;
    return true;
}

void store_tuning(const std::string& key, const Tuning& t) {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tuned[key] = t;
}

RowPlan plan_rows(int mx, int my, size_t elem, int nt_min);

// The three plans of an instance for a tuning (0 entries: defaults: 32-byte
// tiles for the inverse column pass, 16-byte tiles for the forward one, which
// holds more per-mode state and runs more workgroups per CU with narrow tiles).
void apply_plans(Instance& h, const Tuning& t) {
    const Desc& d = h.desc;
    const int nt_min = h.batch_shape ? 256 : 0;
    h.cols = plan_columns(d.my, d.nay, d.hxa, h.elem, nt_min, 32, t.c1);
    h.cols3 = plan_columns(d.my, d.nay, d.hxa, h.elem, nt_min, 16, t.c3);
    h.rows = plan_rows(d.mx, d.my, h.elem, std::max(nt_min, t.rnt));
    const int rt = t.row_transform, ct = t.col_transform;
    if (rt == 1 || rt == 2) h.rows.fl = 0;
    if (rt >= 3 && h.rows.fl > 0) {
        const size_t one = static_cast<size_t>(lp_size(h.rows.R * d.mx)) * h.elem;
        h.rows.nbuf = (rt == 4 && two_buffers_fit(one, h.rows.NT)) ? 2 : 1;
        h.rows.lds = h.rows.nbuf * one;
    }
    // fused column builds: C * My = NT * V and (K3) every kept ky within the
    // first and last thirds of the column
    const int fz_nb = ct == 3 ? 1 : (ct == 4 ? 2 : 0);
    if (fz_nb > 0 && d.my % 3 == 0 && 3 * d.nay <= 2 * d.my) {
        for (ColPlan* p : {&h.cols, &h.cols3}) {
            const int v = 6;
            auto fused_for = [&](int c) { return (c * d.my) % v == 0 && col_fused_built(c * d.my / v, v, d.my); };
            // a tile that leaves the lanes half empty (C * My < NT * 6) takes
            // the columns that fill them when that build exists
            if (p->C * d.my < p->NT * v && (p->NT * v) % d.my == 0) {
                const int c2 = p->NT * v / d.my;
                if ((c2 & (c2 - 1)) == 0 && c2 <= pow2ceil(d.hxa) && col_fused_built(p->NT, v, d.my)) {
                    p->C = c2;
                    p->ntiles = (d.hxa + c2 - 1) / c2;
                }
            }
            // an untuned tile with no fused build takes the widest narrower
            // one that has it (a tuned tile stays as it was measured)
            if ((p == &h.cols ? t.c1 : t.c3) <= 0 && !fused_for(p->C)) {
                for (int c2 = p->C >> 1; c2 >= 1; c2 >>= 1) {
                    if (!fused_for(c2)) continue;
                    p->C = c2;
                    p->ntiles = (d.hxa + c2 - 1) / c2;
                    break;
                }
            }
            if (!fused_for(p->C)) continue;
            const size_t one = static_cast<size_t>(lp_size(p->C * d.my)) * h.elem;
            p->NT = p->C * d.my / v;
            p->V = v;
            p->acc = 2 * v / 3;
            p->fz = d.my;
            p->nbuf = (fz_nb == 2 && two_buffers_fit(one, p->NT)) ? 2 : 1;
            p->lds = p->nbuf * one;
            p->lds3 = std::max(p->nbuf * one, static_cast<size_t>(2 * p->C * d.nay) * h.elem);
        }
    }
    // A fused forward-column tile of one wave whose launch has few workgroups
    // (fewer than half the CUs) runs as four wave groups that split its
    // products: the transforms of a tile and its stage update then take a
    // quarter of the rounds (small grids are bound by that chain, not by HBM).
    const int kFewBlocks = 128;
    {
        RowPlan& r = h.rows;
        const bool few = !h.batch_shape && r.nblocks <= kFewBlocks;
        if (t.k2ws != 1 && r.NT == 64 && r.fl == 0 && r.V == 6 && row_ws_built(4 * r.NT, r.V) && (t.k2ws == 4 || few)) {
            r.ws = 4;
            r.NT = 4 * 64;
            r.lds = 4 * static_cast<size_t>(lp_size(r.R * d.mx)) * h.elem +
                    static_cast<size_t>(std::max(d.n_in, 1)) * r.R * d.mx * (h.elem / 2);   // transforms, physical fields
        }
    }
    {
        ColPlan& p = h.cols3;
        const bool few = !h.batch_shape && p.ntiles * d.ngroups <= kFewBlocks;
        if (t.k3ws != 1 && p.fz > 0 && p.NT == 64 && col_ws_built(4 * p.NT, p.V, p.fz) && (t.k3ws == 4 || few)) {
            const size_t one = static_cast<size_t>(lp_size(p.C * d.my)) * h.elem;
            p.ws = 4;
            p.NT = 4 * 64;
            p.lds3 = std::max(4 * p.nbuf * one, static_cast<size_t>(2 * 4 * p.C * d.nay) * h.elem);
        }
    }
    h.desc.ulg = (t.ulg > 0 && t.ulg <= 12 && d.hxa % (1 << t.ulg) == 0) ? t.ulg : 0;
    h.desc.tlg = (t.tlg >= 0 && t.tlg <= 12 && d.hxa % (1 << t.tlg) == 0) ? t.tlg : -1;
    h.desc.fixed_y = (t.col_transform == 0 || t.col_transform >= 3) ? 1 : 0;
    h.desc.plan_y = make_fft_plan(d.my, t.col_transform == 2 ? 4 : 8);
    h.desc.plan_x = make_fft_plan(d.mx, t.row_transform == 2 ? 4 : 8);
    int npsi_roles = 0;
    for (int i = 0; i < d.n_in; ++i)
        if (d.in_role[i] >= ROLE_PSI0) npsi_roles = std::max(npsi_roles, d.in_role[i] - ROLE_PSI0 + 1);
    h.rows.maxn = npsi_roles == 0 ? 0 : (npsi_roles == 1 ? 1 : (npsi_roles == 2 ? 2 : (npsi_roles <= 4 ? 4 : 8)));
    // a fused row build for this exact field order, where one is built
    h.rows.layout = 0;
    if (h.rows.nbuf > 0 && h.rows.maxn == npsi_roles) {
        for (int lay = 1; lay <= 3 && h.rows.layout == 0; ++lay) {
            if (layout_nin(lay, npsi_roles) != d.n_in || d.n_in == 0) continue;
            bool match = true;
            for (int p = 0; p < d.n_in; ++p) match = match && d.in_role[p] == layout_role(lay, npsi_roles, p);
            if (match && row_layout_built(h.rows.NT, h.rows.V, h.rows.maxn, h.rows.fl, h.rows.nbuf, lay))
                h.rows.layout = lay;
        }
    }
}

// bytes a configuration holds on the device (state x3, work planes, twiddles)
// The parts of an instance's one device allocation, 256-byte aligned: state
// x3, FP64 k = 0 copies x3 (FP32 form), T planes, U planes, x and y twiddle
// tables.
struct SlabParts {
    size_t state, k0, t, u, twx, twy;
    size_t total() const { return 3 * state + 3 * k0 + t + u + twx + twy; }
};
SlabParts slab_parts(const Instance& h) {
    const size_t nf = static_cast<size_t>(h.desc.nfields);
    auto al = [](size_t n) { return (n + 255) & ~static_cast<size_t>(255); };
    SlabParts s;
    s.state = al(nf * h.field_elems * h.elem);
    s.k0 = h.fp32 ? al(nf * sizeof(cx<double>)) : 0;
    s.t = al(std::max<size_t>(1, static_cast<size_t>(h.n_in)) * h.plane_elems * h.elem);
    s.u = al(std::max<size_t>(1, static_cast<size_t>(h.n_out)) * h.plane_elems * h.elem);
    s.twx = al(static_cast<size_t>(fz_tw_size(h.cfg.nx)) * h.elem);
    s.twy = al(static_cast<size_t>(fz_tw_size(h.cfg.ny)) * h.elem);
    return s;
}

size_t footprint_of(const Instance& h) { return slab_parts(h).total(); }

// Lower and plan an instance (no device memory yet).
shhd_status prepare(const shhd_config& cfg, bool batch_shape, std::unique_ptr<Instance>& out,
                    const Tuning* force = nullptr) {
    shhd_status st = validate(cfg);
    if (st != SHHD_OK) return st;
    auto h = std::make_unique<Instance>();
    h->cfg = cfg;
    h->fp32 = cfg.precision == SHHD_PRECISION_FP32;
    h->elem = h->fp32 ? sizeof(cx<float>) : sizeof(cx<double>);
    h->stages = cfg.integrator == SHHD_INTEGRATOR_SRK3 ? 3 : 1;
    st = build_desc(h.get());
    if (st != SHHD_OK) return st;
    const Desc& d = h->desc;
    h->batch_shape = batch_shape;
    h->tune_key = tune_key_of(*h);
    Tuning t = default_tuning();
    if (force) t = *force;
    else lookup_tuning(h->tune_key, t);
    apply_plans(*h, t);
    h->field_elems = static_cast<size_t>(d.hxa) * d.nay;
    h->plane_elems = static_cast<size_t>(d.my) * d.hxa;
    h->bytes = footprint_of(*h);
    out = std::move(h);
    return SHHD_OK;
}

template <typename T>
shhd_status upload_twiddles(Instance* h, std::string* err) {
    std::vector<cx<T>> tx, ty;
    make_twiddles<T>(h->cfg.nx, tx);
    make_twiddles<T>(h->cfg.ny, ty);
    SHHD_HIP(err, hipMemcpy(h->twx, tx.data(), tx.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
    SHHD_HIP(err, hipMemcpy(h->twy, ty.data(), ty.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
    return SHHD_OK;
}

// One device allocation per instance, carved into 256-byte aligned parts:
// state x3 and the FP64 k = 0 copies first (cleared with one memset), then
// the T and U planes and the two twiddle tables.
shhd_status allocate(Instance* h, std::string* err) {
    const SlabParts parts = slab_parts(*h);
    const size_t sb = parts.state, kb = parts.k0, tb = parts.t, ub = parts.u, xb = parts.twx;
    SHHD_HIP(err, hipMalloc(&h->slab, parts.total()));
    char* p = static_cast<char*>(h->slab);
    for (int b = 0; b < 3; ++b, p += sb) h->state[b] = p;
    if (h->fp32)
        for (int b = 0; b < 3; ++b, p += kb) h->k0[b] = reinterpret_cast<cx<double>*>(p);
    SHHD_HIP(err, hipMemset(h->slab, 0, 3 * sb + 3 * kb));
    h->tbuf = p;
    p += tb;
    h->ubuf = p;
    p += ub;
    h->twx = p;
    p += xb;
    h->twy = p;
    return h->fp32 ? upload_twiddles<float>(h, err) : upload_twiddles<double>(h, err);
}

// Reference-layout state -> compact device layout (into buffer h->cur).
shhd_status upload_state(Instance* h, const double* in, size_t n, hipStream_t stream, std::string* err) {
    const shhd_config& c = h->cfg;
    const size_t spec = ref_spec(c);
    const int nf = h->desc.nfields;
    if (n != 2 * spec * static_cast<size_t>(nf)) return set_err(err, SHHD_ERR_INVALID, "state size mismatch");
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
                        return set_err(err, SHHD_ERR_INVALID,
                                       "state has nonzero inactive modes (the 3/2-rule mask must be applied)");
                    continue;
                }
                const int ky = (gy <= c.ny / 2) ? gy : gy - c.ny;
                const int cidx = ky >= 0 ? ky : ky + nay;
                compact[static_cast<size_t>(f) * h->field_elems + static_cast<size_t>(gx) * nay + cidx] = mk(re, im);
            }
        }
    }
    SHHD_HIP(err, hipStreamSynchronize(stream));
    // static fields live in buffer 0, the others in the current buffer
    for (int f = 0; f < nf; ++f) {
        const int b = h->is_static[f] ? 0 : h->cur;
        const size_t off = static_cast<size_t>(f) * h->field_elems;
        if (h->fp32) {
            std::vector<cx<float>> cf(h->field_elems);
            for (size_t i = 0; i < cf.size(); ++i)
                cf[i] = mk(static_cast<float>(compact[off + i].re), static_cast<float>(compact[off + i].im));
            SHHD_HIP(err, hipMemcpy(static_cast<cx<float>*>(h->state[b]) + off, cf.data(), cf.size() * sizeof(cx<float>),
                                    hipMemcpyHostToDevice));
            SHHD_HIP(err, hipMemcpy(h->k0[b] + f, &compact[off], sizeof(cx<double>), hipMemcpyHostToDevice));
        } else {
            SHHD_HIP(err, hipMemcpy(static_cast<cx<double>*>(h->state[b]) + off, compact.data() + off,
                                    h->field_elems * sizeof(cx<double>), hipMemcpyHostToDevice));
        }
    }
    // reference density rho_hat(0,0)/(Mx*My) (src/fcalculator_workspace.cc:41-66)
    h->desc.rho0 = compact[0].re / (static_cast<double>(c.nx) * static_cast<double>(c.ny));
    return SHHD_OK;
}

// Compact per-field columns -> reference layout (active modes only; the rest
// of `out` is left as the caller cleared it).
void compact_to_reference(const Instance& h, const cx<double>* compact, int nf, double* out) {
    const shhd_config& c = h.cfg;
    const size_t spec = ref_spec(c);
    const int hx = c.nx / 2 + 1;
    const int nay = c.ny_active;
    for (int f = 0; f < nf; ++f) {
        double* dst = out + 2 * spec * static_cast<size_t>(f);
        for (int gy = 0; gy < c.ny; ++gy) {
            for (int gx = 0; gx < c.nx_active / 2; ++gx) {
                if (!ref_active(c, gx, gy)) continue;
                const int ky = (gy <= c.ny / 2) ? gy : gy - c.ny;
                const int cidx = ky >= 0 ? ky : ky + nay;
                const cx<double> v = compact[static_cast<size_t>(f) * h.field_elems + static_cast<size_t>(gx) * nay + cidx];
                const size_t i = static_cast<size_t>(gy) * hx + gx;
                dst[2 * i] = v.re;
                dst[2 * i + 1] = v.im;
            }
        }
    }
}

shhd_status download_state(const Instance* h, double* out, size_t n, hipStream_t stream, std::string* err) {
    const shhd_config& c = h->cfg;
    const size_t spec = ref_spec(c);
    const int nf = h->desc.nfields;
    if (n != 2 * spec * static_cast<size_t>(nf)) return set_err(err, SHHD_ERR_INVALID, "state size mismatch");
    SHHD_HIP(err, hipStreamSynchronize(stream));
    std::vector<cx<double>> compact(static_cast<size_t>(nf) * h->field_elems);
    for (int f = 0; f < nf; ++f) {
        const int b = h->is_static[f] ? 0 : h->cur;
        const size_t off = static_cast<size_t>(f) * h->field_elems;
        if (h->fp32) {
            std::vector<cx<float>> cf(h->field_elems);
            SHHD_HIP(err, hipMemcpy(cf.data(), static_cast<cx<float>*>(h->state[b]) + off, cf.size() * sizeof(cx<float>),
                                    hipMemcpyDeviceToHost));
            for (size_t i = 0; i < cf.size(); ++i) compact[off + i] = mk<double>(cf[i].re, cf[i].im);
            cx<double> k0v;
            SHHD_HIP(err, hipMemcpy(&k0v, h->k0[b] + f, sizeof(cx<double>), hipMemcpyDeviceToHost));
            compact[off] = k0v;
        } else {
            SHHD_HIP(err, hipMemcpy(compact.data() + off, static_cast<cx<double>*>(h->state[b]) + off,
                                    h->field_elems * sizeof(cx<double>), hipMemcpyDeviceToHost));
        }
    }
    std::memset(out, 0, sizeof(double) * n);
    compact_to_reference(*h, compact.data(), nf, out);
    return SHHD_OK;
}

// ---------------------------------------------------------------------------
// engine: instances grouped into launch classes
// ---------------------------------------------------------------------------
struct ClassKey {
    bool fp32;
    int stages, cNT, cV, c3NT, c3V, acc, rNT, rV, maxn, rfl;
    int cfz, cnb, c3fz, c3nb, rnb, rlay;   // fused builds (length, LDS buffers, field layout)
    int c3ws, rws;                         // forward-column and row wave groups
    bool operator<(const ClassKey& o) const {
        return std::tie(fp32, stages, cNT, cV, c3NT, c3V, acc, rNT, rV, maxn, rfl, cfz, cnb, c3fz, c3nb, rnb, rlay, c3ws,
                        rws) <
               std::tie(o.fp32, o.stages, o.cNT, o.cV, o.c3NT, o.c3V, o.acc, o.rNT, o.rV, o.maxn, o.rfl, o.cfz, o.cnb,
                        o.c3fz, o.c3nb, o.rnb, o.rlay, o.c3ws, o.rws);
    }
};

struct LaunchClass {
    ClassKey key{};
    std::vector<int> members;        // slots
    void* d_inst = nullptr;          // InstDev<T>[n]
    void* d_ptrs = nullptr;          // StagePtrs<T>[3 * stages][n]
    int* d_map = nullptr;            // block maps of K1, K2, K3
    void* d_static = nullptr;        // StaticEntry<T>[nstatic]
    int nstatic = 0;
    void* d_meas = nullptr;          // MeasDev<T>[n] when any member measures
    bool has_probe = false, has_sk = false;
    int sk_blocks = 0;
    int nblk[3] = {0, 0, 0};
    BlockMap map[3] = {};
    size_t lds[3] = {0, 0, 0};

    void release() {
        if (d_inst) (void)hipFree(d_inst);
        if (d_ptrs) (void)hipFree(d_ptrs);
        if (d_map) (void)hipFree(d_map);
        if (d_static) (void)hipFree(d_static);
        if (d_meas) (void)hipFree(d_meas);
        d_inst = d_ptrs = d_static = d_meas = nullptr;
        d_map = nullptr;
    }
};

struct Engine {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    bool batch_shape = false;     // plan small grids onto the shared 256-lane shape
    bool use_graph = false;
    // batch: an instance whose density reached zero stays stopped until it is
    // retired; single handle: every step call starts from a cleared flag
    bool sticky_failures = false;
    size_t budget = 0, used = 0;
    std::vector<std::unique_ptr<Instance>> slots;
    uint64_t* d_steps = nullptr;
    int* d_flags = nullptr;
    int* d_active = nullptr;      // 1 for the slots in a launch class (they step)
    std::vector<LaunchClass> classes;
    bool dirty = true;
    // steps taken since the launch tables were built, mod 3: the tables hold
    // the three buffer rotations from the one at build time, so a step at
    // rotation r uses entry r and no call needs a rebuild for its step count
    int rot = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;

    // Launch classes run on up to kLaneStreams streams (this one and side
    // streams of default priority): instances of different classes share no
    // buffers, so a step forks from `stream`, each class runs its stages on
    // its own stream, and the step joins back before the step counters
    // advance.  Small grids then fill the GPU beside large ones instead of
    // waiting for them.  `ls` is the stream the class being issued uses.
    static constexpr int kLaneStreams = 4;
    hipStream_t side[kLaneStreams - 1] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[kLaneStreams - 1] = {};
    hipStream_t ls = nullptr;

    bool timing = false;
    double time_ms[kClasses] = {0, 0, 0};
    long long launches[kClasses] = {0, 0, 0};
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;

    ~Engine() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (int i = 0; i < kLaneStreams - 1; ++i)
            if (side[i]) (void)hipStreamSynchronize(side[i]);
        drop_graph();
        for (auto& c : classes) c.release();
        slots.clear();
        if (d_steps) (void)hipFree(d_steps);
        if (d_flags) (void)hipFree(d_flags);
        if (d_active) (void)hipFree(d_active);
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        for (auto& u : ev_used) {
            (void)hipEventDestroy(u.second.first);
            (void)hipEventDestroy(u.second.second);
        }
        for (int i = 0; i < kLaneStreams - 1; ++i) {
            if (side[i]) (void)hipStreamDestroy(side[i]);
            if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
        }
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // streams the classes of a step are spread over (1: all on `stream`)
    int lane_streams() const {
        if (timing || classes.size() < 2) return 1;
        return std::min<int>(kLaneStreams, static_cast<int>(classes.size()));
    }
    shhd_status ensure_side_streams() {
        if (fork_ev) return SHHD_OK;
        for (int i = 0; i < kLaneStreams - 1; ++i) {
            SHHD_HIP(&err, hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking));
            SHHD_HIP(&err, hipEventCreateWithFlags(&join_ev[i], hipEventDisableTiming));
        }
        SHHD_HIP(&err, hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
        return SHHD_OK;
    }
    hipStream_t class_stream(size_t c, int nls) const { return (c % nls) == 0 ? stream : side[c % nls - 1]; }

    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }

    shhd_status init(size_t budget_bytes, bool batch, bool graphs, bool sticky = false) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
            return set_err(&err, SHHD_ERR_HIP, "no HIP device available");
        SHHD_HIP(&err, hipGetDevice(&device));
        SHHD_HIP(&err, hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        SHHD_HIP(&err, hipMalloc(&d_steps, kMaxSlots * sizeof(uint64_t)));
        SHHD_HIP(&err, hipMalloc(&d_flags, kMaxSlots * sizeof(int)));
        SHHD_HIP(&err, hipMemset(d_steps, 0, kMaxSlots * sizeof(uint64_t)));
        SHHD_HIP(&err, hipMemset(d_flags, 0, kMaxSlots * sizeof(int)));
        SHHD_HIP(&err, hipMalloc(&d_active, kMaxSlots * sizeof(int)));
        SHHD_HIP(&err, hipMemset(d_active, 0, kMaxSlots * sizeof(int)));
        budget = budget_bytes;
        batch_shape = batch;
        use_graph = graphs;
        sticky_failures = sticky;
        slots.resize(kMaxSlots);
        return SHHD_OK;
    }

    int active_count() const {
        int n = 0;
        for (auto& s : slots) n += s ? 1 : 0;
        return n;
    }

    shhd_status admit(const shhd_config& cfg, const double* state, size_t n, int* slot_out) {
        std::unique_ptr<Instance> h;
        shhd_status st = prepare(cfg, batch_shape, h);
        if (st != SHHD_OK) return set_err(&err, st, g_create_error);
        if (budget && used + h->bytes > budget)
            return set_err(&err, SHHD_ERR_FULL, "device budget exhausted: instance needs " + std::to_string(h->bytes) +
                                                     " bytes, " + std::to_string(budget - used) + " left");
        int slot = -1;
        for (int i = 0; i < kMaxSlots; ++i)
            if (!slots[static_cast<size_t>(i)]) { slot = i; break; }
        if (slot < 0) return set_err(&err, SHHD_ERR_FULL, "no free instance slot");
        st = allocate(h.get(), &err);
        if (st != SHHD_OK) return st;
        if (state) {
            st = upload_state(h.get(), state, n, stream, &err);
            if (st != SHHD_OK) return st;
        }
        const uint64_t zero = 0;
        const int ok = 0;
        SHHD_HIP(&err, hipMemcpy(d_steps + slot, &zero, sizeof(zero), hipMemcpyHostToDevice));
        SHHD_HIP(&err, hipMemcpy(d_flags + slot, &ok, sizeof(ok), hipMemcpyHostToDevice));
        used += h->bytes;
        slots[static_cast<size_t>(slot)] = std::move(h);
        dirty = true;
        *slot_out = slot;
        return SHHD_OK;
    }

    shhd_status retire(int slot) {
        if (slot < 0 || slot >= kMaxSlots || !slots[static_cast<size_t>(slot)])
            return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        used -= slots[static_cast<size_t>(slot)]->bytes;
        slots[static_cast<size_t>(slot)].reset();
        dirty = true;
        return SHHD_OK;
    }

    Instance* at(int slot) {
        if (slot < 0 || slot >= kMaxSlots) return nullptr;
        return slots[static_cast<size_t>(slot)].get();
    }

    // new launch shapes for an instance (the buffers do not depend on them)
    shhd_status replan(int slot, const Tuning& t) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        apply_plans(*h, t);
        dirty = true;
        return SHHD_OK;
    }

    // FluidSolver::set_flux_request (src/fluid_solver.h:101-103).  The buffer
    // is allocated on the first request; a changed request clears every slot
    // before the next step, as FluxBuffer::begin_step does.
    shhd_status set_flux_request(int slot, int density, int order_parameter, int momentum) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        const int req = (density ? FLUX_DENSITY : 0) | (order_parameter ? FLUX_ORDER_PARAMETER : 0) |
                        (momentum ? FLUX_MOMENTUM : 0);
        if (req && !h->flux) {
            const size_t fb = static_cast<size_t>(h->flux_slots()) * h->field_elems * h->elem;
            if (budget && used + fb > budget)
                return set_err(&err, SHHD_ERR_FULL, "device budget exhausted: flux buffer needs " + std::to_string(fb) + " bytes");
            SHHD_HIP(&err, hipStreamSynchronize(stream));
            SHHD_HIP(&err, hipMalloc(&h->flux, fb));
            SHHD_HIP(&err, hipMemset(h->flux, 0, fb));
            h->bytes += fb;
            used += fb;
        }
        if (req != h->flux_req) {
            if (req) h->flux_clear = true;
            h->flux_req = req;
            dirty = true;
        }
        return SHHD_OK;
    }

    // Mode probes: after every step the values of the chosen (state field or
    // flux slot, reference mode) pairs are appended to a ring of `capacity`
    // FP64 rows.  nprobes = 0 removes them.
    shhd_status set_probes(int slot, const shhd_probe* probes, int nprobes, int capacity) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        if (nprobes < 0 || nprobes > kMaxProbes)
            return set_err(&err, SHHD_ERR_INVALID, "number of probes must be in [0, " + std::to_string(kMaxProbes) + "]");
        if (nprobes > 0 && (!probes || capacity <= 0))
            return set_err(&err, SHHD_ERR_INVALID, "probes need a list and a positive capacity");
        const shhd_config& c = h->cfg;
        std::vector<Probe> list;
        for (int p = 0; p < nprobes; ++p) {
            const shhd_probe& q = probes[p];
            const int limit = q.source == 0 ? h->desc.nfields : h->flux_slots();
            if ((q.source != 0 && q.source != 1) || q.index < 0 || q.index >= limit)
                return set_err(&err, SHHD_ERR_INVALID, "probe " + std::to_string(p) + ": invalid source or index");
            if (q.gx < 0 || q.gx >= c.nx / 2 + 1 || q.gy < 0 || q.gy >= c.ny || !ref_active(c, q.gx, q.gy))
                return set_err(&err, SHHD_ERR_INVALID, "probe " + std::to_string(p) + ": mode is outside the active spectral mask");
            const int ky = (q.gy <= c.ny / 2) ? q.gy : q.gy - c.ny;
            const int cidx = ky >= 0 ? ky : ky + c.ny_active;
            list.push_back({q.source, q.index, q.gx * c.ny_active + cidx});
        }
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        used -= h->probe_bytes;
        h->bytes -= h->probe_bytes;
        h->probe_bytes = 0;
        h->free_probes();
        dirty = true;
        if (nprobes == 0) return SHHD_OK;
        const size_t ring = static_cast<size_t>(capacity) * static_cast<size_t>(nprobes) * sizeof(cx<double>);
        const size_t pb = ring + list.size() * sizeof(Probe) + sizeof(int);
        if (budget && used + pb > budget)
            return set_err(&err, SHHD_ERR_FULL, "device budget exhausted: probe ring needs " + std::to_string(pb) + " bytes");
        SHHD_HIP(&err, hipMalloc(&h->d_probes, list.size() * sizeof(Probe)));
        SHHD_HIP(&err, hipMemcpy(h->d_probes, list.data(), list.size() * sizeof(Probe), hipMemcpyHostToDevice));
        SHHD_HIP(&err, hipMalloc(&h->d_ring, ring));
        SHHD_HIP(&err, hipMalloc(&h->d_rows, sizeof(int)));
        SHHD_HIP(&err, hipMemset(h->d_rows, 0, sizeof(int)));
        h->probes = list;
        h->capacity = capacity;
        h->probe_rows = 0;
        h->probe_bytes = pb;
        h->bytes += pb;
        used += pb;
        return SHHD_OK;
    }

    // The rows recorded since the last read, oldest first, as
    // rows x nprobes FP64 complex; the ring is emptied.
    shhd_status read_probes(int slot, double* out, size_t n, int* nrows) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        if (h->probes.empty()) return set_err(&err, SHHD_ERR_INVALID, "no probes are set");
        const size_t need = 2 * static_cast<size_t>(h->probe_rows) * h->probes.size();
        if (n < need) return set_err(&err, SHHD_ERR_INVALID, "probe output too small: " + std::to_string(need) + " doubles needed");
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        int rows = 0;
        SHHD_HIP(&err, hipMemcpy(&rows, h->d_rows, sizeof(int), hipMemcpyDeviceToHost));
        if (rows != h->probe_rows) return set_err(&err, SHHD_ERR_RUNTIME, "probe ring row count mismatch");
        if (need) SHHD_HIP(&err, hipMemcpy(out, h->d_ring, need * sizeof(double), hipMemcpyDeviceToHost));
        SHHD_HIP(&err, hipMemset(h->d_rows, 0, sizeof(int)));
        if (nrows) *nrows = rows;
        h->probe_rows = 0;
        return SHHD_OK;
    }

    // Structure factor: every `nevery` steps from now on, |f_hat(k)|^2 of each
    // field in field_mask (bit f = state field f) is added to an FP64 sum per
    // active mode.  field_mask = 0 removes it; calling again restarts it.
    shhd_status set_structure_factor(int slot, unsigned field_mask, int nevery) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        if (field_mask >> h->desc.nfields)
            return set_err(&err, SHHD_ERR_INVALID, "structure-factor field mask names a field the state does not have");
        if (field_mask && nevery <= 0) return set_err(&err, SHHD_ERR_INVALID, "structure-factor interval must be positive");
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        used -= h->sk_bytes;
        h->bytes -= h->sk_bytes;
        h->sk_bytes = 0;
        h->free_sk();
        dirty = true;
        if (!field_mask) return SHHD_OK;
        std::vector<int> fields;
        for (int f = 0; f < h->desc.nfields; ++f)
            if ((field_mask >> f) & 1u) fields.push_back(f);
        const size_t sb = fields.size() * h->field_elems * sizeof(double) + sizeof(unsigned long long);
        if (budget && used + sb > budget)
            return set_err(&err, SHHD_ERR_FULL, "device budget exhausted: structure factor needs " + std::to_string(sb) + " bytes");
        SHHD_HIP(&err, hipMalloc(&h->d_sk, fields.size() * h->field_elems * sizeof(double)));
        SHHD_HIP(&err, hipMemset(h->d_sk, 0, fields.size() * h->field_elems * sizeof(double)));
        SHHD_HIP(&err, hipMalloc(&h->d_sk_count, sizeof(unsigned long long)));
        SHHD_HIP(&err, hipMemset(h->d_sk_count, 0, sizeof(unsigned long long)));
        h->sk_fields = fields;
        h->sk_nevery = nevery;
        h->sk_base = h->step_index;
        h->sk_bytes = sb;
        h->bytes += sb;
        used += sb;
        return SHHD_OK;
    }

    // The sums in the reference layout, one real[(Mx/2+1) * My] per selected
    // field in field order (inactive modes 0), and the number of samples.
    shhd_status get_structure_factor(int slot, double* out, size_t n, long long* nsamples) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        if (h->sk_fields.empty()) return set_err(&err, SHHD_ERR_INVALID, "no structure factor is accumulated");
        const shhd_config& c = h->cfg;
        const size_t spec = ref_spec(c);
        const size_t nf = h->sk_fields.size();
        if (n != nf * spec) return set_err(&err, SHHD_ERR_INVALID, "structure-factor size mismatch");
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        std::vector<double> compact(nf * h->field_elems);
        SHHD_HIP(&err, hipMemcpy(compact.data(), h->d_sk, compact.size() * sizeof(double), hipMemcpyDeviceToHost));
        unsigned long long cnt = 0;
        SHHD_HIP(&err, hipMemcpy(&cnt, h->d_sk_count, sizeof(cnt), hipMemcpyDeviceToHost));
        std::memset(out, 0, sizeof(double) * n);
        const int hx = c.nx / 2 + 1;
        const int nay = c.ny_active;
        for (size_t q = 0; q < nf; ++q) {
            for (int gy = 0; gy < c.ny; ++gy) {
                for (int gx = 0; gx < c.nx_active / 2; ++gx) {
                    if (!ref_active(c, gx, gy)) continue;
                    const int ky = (gy <= c.ny / 2) ? gy : gy - c.ny;
                    const int cidx = ky >= 0 ? ky : ky + nay;
                    out[q * spec + static_cast<size_t>(gy) * hx + gx] =
                        compact[q * h->field_elems + static_cast<size_t>(gx) * nay + cidx];
                }
            }
        }
        if (nsamples) *nsamples = static_cast<long long>(cnt);
        return SHHD_OK;
    }

    // `set <target> <style>` evaluated on the device into the resident state
    // (src/initial_condition_*.cc); argument checks and messages follow the
    // styles' parsers and apply() functions.
    // no_alloc: the normalization partials go to `scratch` instead of a
    // per-call hipMalloc (nothing on this path allocates device memory)
    shhd_status apply_initial_condition(int slot, const shhd_initial_condition* icp, bool no_alloc = false) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        if (!icp) return set_err(&err, SHHD_ERR_INVALID, "null initial condition");
        const shhd_initial_condition ic = *icp;
        const shhd_config& c = h->cfg;
        const int N = h->desc.N;
        int f = -1;
        switch (ic.kind) {
            case SHHD_IC_EQUILIBRIUM_DENSITY:
                if (ic.kBT < 0.0) return set_err(&err, SHHD_ERR_INVALID, "set density equilibrium/gaussian requires kBT >= 0.");
                if (ic.value <= 0.0) return set_err(&err, SHHD_ERR_INVALID, "set density equilibrium/gaussian requires rho0 > 0.");
                if (ic.pressure_coefficient <= 0.0)
                    return set_err(&err, SHHD_ERR_INVALID, "set density equilibrium/gaussian requires cT > 0.");
                f = 0;
                break;
            case SHHD_IC_EQUILIBRIUM_ORDER_PARAMETER:
                if (ic.component < 0 || ic.component >= N)
                    return set_err(&err, SHHD_ERR_INVALID, "order-parameter index out of range");
                if (ic.kBT < 0.0) return set_err(&err, SHHD_ERR_INVALID, "set order_parameter equilibrium/gaussian requires kBT >= 0.");
                f = 1 + ic.component;
                break;
            case SHHD_IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE:
            case SHHD_IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE:
            case SHHD_IC_RANDOM_VORTICITY:
            case SHHD_IC_SPECTRUM_K4:
                if (ic.component != 0 && ic.component != 1)
                    return set_err(&err, SHHD_ERR_INVALID, "momentum direction must be 0 or 1");
                if ((ic.kind == SHHD_IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE ||
                     ic.kind == SHHD_IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE) && ic.kBT < 0.0)
                    return set_err(&err, SHHD_ERR_INVALID, std::string("set momentum ") +
                                   (ic.kind == SHHD_IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE
                                        ? "equilibrium/gaussian/compressible" : "equilibrium/gaussian/incompressible") +
                                   " requires kBT >= 0.");
                if (ic.kind == SHHD_IC_RANDOM_VORTICITY || ic.kind == SHHD_IC_SPECTRUM_K4) {
                    const char* nm = ic.kind == SHHD_IC_RANDOM_VORTICITY ? "random_vorticity" : "spectrum/k4";
                    if (ic.n0 <= 0.0) return set_err(&err, SHHD_ERR_INVALID, std::string("set momentum ") + nm + " requires positive n0.");
                    if (ic.kind == SHHD_IC_RANDOM_VORTICITY && ic.sigma <= 0.0)
                        return set_err(&err, SHHD_ERR_INVALID, "set momentum random_vorticity requires positive sigma.");
                    if (ic.urms <= 0.0) return set_err(&err, SHHD_ERR_INVALID, std::string("set momentum ") + nm + " requires positive urms.");
                }
                f = N + 1 + ic.component;
                break;
            default:
                return set_err(&err, SHHD_ERR_INVALID, "unknown initial-condition kind");
        }
        IcDev d{};
        d.ic = ic;
        d.nay = c.ny_active;
        d.hxa = c.nx_active / 2;
        d.mx = c.nx;
        d.my = c.ny;
        d.lx = c.lx;
        d.ly = c.ly;
        d.grid = static_cast<double>(c.nx) * static_cast<double>(c.ny);
        d.cell = c.lx * c.ly / d.grid;
        const double two_pi = 2.0 * 3.14159265358979323846;
        if (ic.kind == SHHD_IC_EQUILIBRIUM_ORDER_PARAMETER && ic.kBT != 0.0) {
            for (int gx = 0; gx < d.hxa; ++gx) {
                for (int ky = -(d.nay / 2 - 1); ky < d.nay / 2; ++ky) {
                    const double kx = two_pi * gx / c.lx, kyv = two_pi * ky / c.ly;
                    const double k2 = kx * kx + kyv * kyv;
                    if (k2 == 0.0) continue;
                    if (ic.k0 + ic.k2 * k2 + ic.k4 * k2 * k2 <= 0.0)
                        return set_err(&err, SHHD_ERR_INVALID,
                                       "set order_parameter equilibrium/gaussian requires positive linear chemical-potential "
                                       "coefficient for every active nonzero mode.");
                }
            }
        }
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        const int b0 = h->is_static[0] ? 0 : h->cur;
        double rho_hat0 = 0.0;
        if (h->fp32) {
            cx<double> v;
            SHHD_HIP(&err, hipMemcpy(&v, h->k0[b0], sizeof(v), hipMemcpyDeviceToHost));
            rho_hat0 = v.re;
        } else {
            cx<double> v;
            SHHD_HIP(&err, hipMemcpy(&v, h->state[b0], sizeof(v), hipMemcpyDeviceToHost));
            rho_hat0 = v.re;
        }
        d.rho0 = rho_hat0 / d.grid;
        const size_t total = static_cast<size_t>(d.hxa) * d.nay;
        const int nblk = static_cast<int>(std::min<size_t>(kIcNormBlocks, (total + 255) / 256));
        if (ic.kind == SHHD_IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE || ic.kind == SHHD_IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE) {
            if (ic.kBT != 0.0 && d.rho0 <= 0.0)
                return set_err(&err, SHHD_ERR_INVALID, std::string("set momentum ") +
                               (ic.kind == SHHD_IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE
                                    ? "equilibrium/gaussian/compressible" : "equilibrium/gaussian/incompressible") +
                               " requires positive reference density.");
        }
        if (ic.kind == SHHD_IC_RANDOM_VORTICITY || ic.kind == SHHD_IC_SPECTRUM_K4) {
            // one partial per workgroup, summed on the host in workgroup order
            if (nblk > kIcNormBlocks)
                return set_err(&err, SHHD_ERR_INVALID, "initial-condition normalization needs " + std::to_string(nblk) +
                                                           " partial sums, more than the " +
                                                           std::to_string(kIcNormBlocks) + " the scratch holds");
            double ps[kIcNormBlocks];
            double* partials = nullptr;
            if (no_alloc) {
                const shhd_status sa = scratch_address(device, &partials, &err);
                if (sa != SHHD_OK) return sa;
            } else {
                SHHD_HIP(&err, hipMalloc(&partials, nblk * sizeof(double)));
            }
            hipLaunchKernelGGL(k_ic_norm, dim3(nblk), dim3(256), 0, stream, d, partials);
            const hipError_t e1 = hipGetLastError();
            const hipError_t e2 = hipMemcpyAsync(ps, partials, nblk * sizeof(double), hipMemcpyDeviceToHost, stream);
            const hipError_t e3 = hipStreamSynchronize(stream);
            if (!no_alloc) (void)hipFree(partials);
            if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
                return set_err(&err, SHHD_ERR_HIP, "initial-condition normalization failed");
            double norm = 0.0;
            for (int q = 0; q < nblk; ++q) norm += ps[q];
            if (norm == 0.0) return set_err(&err, SHHD_ERR_INVALID, "set momentum random_vorticity generated zero velocity field.");
            if (d.rho0 == 0.0) return set_err(&err, SHHD_ERR_INVALID, "set momentum random_vorticity requires nonzero reference density.");
            d.scale = ic.urms * d.grid / std::sqrt(norm);
        }
        const int b = h->is_static[f] ? 0 : h->cur;
        const size_t off = static_cast<size_t>(f) * h->field_elems;
        if (h->fp32)
            hipLaunchKernelGGL(k_initial_condition<float>, dim3(nblk), dim3(256), 0, stream, d,
                               static_cast<cx<float>*>(h->state[b]) + off);
        else
            hipLaunchKernelGGL(k_initial_condition<double>, dim3(nblk), dim3(256), 0, stream, d,
                               static_cast<cx<double>*>(h->state[b]) + off);
        SHHD_HIP(&err, hipGetLastError());
        double k0re = 0.0;
        if (ic.kind == SHHD_IC_EQUILIBRIUM_DENSITY || ic.kind == SHHD_IC_EQUILIBRIUM_ORDER_PARAMETER) k0re = ic.value * d.grid;
        if (h->fp32) {
            const cx<double> v = mk(k0re, 0.0);
            SHHD_HIP(&err, hipMemcpyAsync(h->k0[b] + f, &v, sizeof(v), hipMemcpyHostToDevice, stream));
        }
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        if (f == 0) {
            h->desc.rho0 = k0re / d.grid;   // as upload_state derives it
            dirty = true;
        }
        return SHHD_OK;
    }

    // All fields in physical space on the padded grid, as the reference's
    // PhysicalStateBuffer after FourierTransform2D::backward_many
    // (src/measure_workspace.h:27-35): out[f][gy][gx], n = nfields*Mx*My.
    // K1 over every field into scratch planes, then the row half.
    template <typename T>
    shhd_status physical_fields(Instance& h, double* out) {
        const int nf = h.desc.nfields;
        const shhd_config& c = h.cfg;
        const size_t nphys = static_cast<size_t>(c.nx) * c.ny;
        InstDev<T> I{};
        I.d = h.desc;
        I.d.n_in = nf;
        for (int f = 0; f < nf; ++f) I.d.in_field[f] = f;
        I.twx = static_cast<const cx<T>*>(h.twx);
        I.twy = static_cast<const cx<T>*>(h.twy);
        I.C = h.cols.C;
        I.R = h.rows.R;
        I.ntiles = h.cols.ntiles;
        const StagePtrs<T> P = stage_ptrs<T>(h, h.cur, h.cur, h.cur);
        const int nblk1 = h.cols.ntiles * nf;
        std::vector<int> map(static_cast<size_t>(nblk1) + 1, 0);   // every block -> instance 0; first[0] = 0
        void* d_t = nullptr;
        void* d_out = nullptr;
        void* d_tab = nullptr;
        shhd_status st = SHHD_OK;
        auto fail = [&](hipError_t e, const char* what) {
            st = set_err(&err, SHHD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
        };
        hipError_t e = hipMalloc(&d_t, static_cast<size_t>(nf) * h.plane_elems * sizeof(cx<T>));
        if (e != hipSuccess) fail(e, "hipMalloc");
        if (st == SHHD_OK && (e = hipMalloc(&d_out, nf * nphys * sizeof(T))) != hipSuccess) fail(e, "hipMalloc");
        const size_t tab_bytes = sizeof(InstDev<T>) + sizeof(StagePtrs<T>) + map.size() * sizeof(int);
        if (st == SHHD_OK && (e = hipMalloc(&d_tab, tab_bytes)) != hipSuccess) fail(e, "hipMalloc");
        if (st == SHHD_OK) {
            I.tbuf = static_cast<cx<T>*>(d_t);
            unsigned char* base = static_cast<unsigned char*>(d_tab);
            InstDev<T>* dI = reinterpret_cast<InstDev<T>*>(base);
            StagePtrs<T>* dP = reinterpret_cast<StagePtrs<T>*>(base + sizeof(InstDev<T>));
            int* dmap = reinterpret_cast<int*>(base + sizeof(InstDev<T>) + sizeof(StagePtrs<T>));
            BlockMap bm{dmap, dmap + nblk1};
            if ((e = hipMemcpy(dI, &I, sizeof(I), hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(dP, &P, sizeof(P), hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(dmap, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess) {
                fail(e, "hipMemcpy");
            } else {
                K1Fn<T> k1 = pick_k1<T>(h.cols.NT, h.cols.V, h.cols.fz, h.cols.nbuf);
                KPFn<T> kp = pick_kp<T>(h.rows.NT, h.rows.V);
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(k1), hipFuncAttributeMaxDynamicSharedMemorySize,
                                          static_cast<int>(h.cols.lds));
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kp), hipFuncAttributeMaxDynamicSharedMemorySize,
                                          static_cast<int>(h.rows.lds));
                hipLaunchKernelGGL(k1, dim3(nblk1), dim3(h.cols.NT), h.cols.lds, stream, dI, dP, bm);
                hipLaunchKernelGGL(kp, dim3(h.rows.nblocks), dim3(h.rows.NT), h.rows.lds, stream, dI,
                                   static_cast<T*>(d_out), nf);
                if ((e = hipGetLastError()) != hipSuccess || (e = hipStreamSynchronize(stream)) != hipSuccess) {
                    fail(e, "physical transform");
                } else {
                    std::vector<T> host(nf * nphys);
                    if ((e = hipMemcpy(host.data(), d_out, host.size() * sizeof(T), hipMemcpyDeviceToHost)) != hipSuccess)
                        fail(e, "hipMemcpy");
                    else
                        for (size_t i = 0; i < host.size(); ++i) out[i] = static_cast<double>(host[i]);
                }
            }
        }
        if (d_t) (void)hipFree(d_t);
        if (d_out) (void)hipFree(d_out);
        if (d_tab) (void)hipFree(d_tab);
        return st;
    }

    shhd_status get_physical(int slot, double* out, size_t n) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        const shhd_config& c = h->cfg;
        if (n != static_cast<size_t>(h->desc.nfields) * c.nx * c.ny)
            return set_err(&err, SHHD_ERR_INVALID, "physical size mismatch");
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        return h->fp32 ? physical_fields<float>(*h, out) : physical_fields<double>(*h, out);
    }

    // FluidSolver::flux_buffer (src/fluid_solver.h:105-107) in the reference
    // layout: slots of complex[(Mx/2+1) * My], inactive modes 0.
    shhd_status get_flux(int slot, double* out, size_t n) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        const shhd_config& c = h->cfg;
        const size_t spec = ref_spec(c);
        const int ns = h->flux_slots();
        if (n != 2 * spec * static_cast<size_t>(ns)) return set_err(&err, SHHD_ERR_INVALID, "flux size mismatch");
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        std::memset(out, 0, sizeof(double) * n);
        if (!h->flux) return SHHD_OK;
        std::vector<cx<double>> compact(static_cast<size_t>(ns) * h->field_elems);
        if (h->fp32) {
            std::vector<cx<float>> cf(compact.size());
            SHHD_HIP(&err, hipMemcpy(cf.data(), h->flux, cf.size() * sizeof(cx<float>), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < cf.size(); ++i) compact[i] = mk<double>(cf[i].re, cf[i].im);
        } else {
            SHHD_HIP(&err, hipMemcpy(compact.data(), h->flux, compact.size() * sizeof(cx<double>), hipMemcpyDeviceToHost));
        }
        compact_to_reference(*h, compact.data(), ns, out);
        return SHHD_OK;
    }

    template <typename T>
    StagePtrs<T> stage_ptrs(const Instance& h, int next, int cur, int base) const {
        StagePtrs<T> p{};
        for (int f = 0; f < h.desc.nfields; ++f) {
            auto at = [&](int b) {
                const int bb = h.is_static[f] ? 0 : b;
                return static_cast<cx<T>*>(h.state[bb]) + static_cast<size_t>(f) * h.field_elems;
            };
            p.cur[f] = at(cur);
            p.base[f] = at(base);
            p.next[f] = at(next);
        }
        if (h.fp32) {
            p.cur_k0 = h.k0[cur];
            p.base_k0 = h.k0[base];
            p.next_k0 = h.k0[next];
        }
        return p;
    }

    template <typename T>
    shhd_status build_class_tables(LaunchClass& lc) {
        const int n = static_cast<int>(lc.members.size());
        const int stages = lc.key.stages;
        std::vector<InstDev<T>> inst(static_cast<size_t>(n));
        std::vector<StagePtrs<T>> ptrs(static_cast<size_t>(3 * stages * n));
        std::vector<int> maps[3];
        std::vector<int> firsts[3];
        std::vector<StaticEntry<T>> statics;
        for (int k = 0; k < 3; ++k) firsts[k].resize(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            Instance& h = *slots[static_cast<size_t>(lc.members[static_cast<size_t>(i)])];
            InstDev<T>& I = inst[static_cast<size_t>(i)];
            I.d = h.desc;
            I.twx = static_cast<const cx<T>*>(h.twx);
            I.twy = static_cast<const cx<T>*>(h.twy);
            I.tbuf = static_cast<cx<T>*>(h.tbuf);
            I.ubuf = static_cast<cx<T>*>(h.ubuf);
            I.flag = d_flags + lc.members[static_cast<size_t>(i)];
            I.step = d_steps + lc.members[static_cast<size_t>(i)];
            I.flux = h.flux_req ? static_cast<cx<T>*>(h.flux) : nullptr;
            I.flux_req = h.flux_req;
            I.C = h.cols.C;
            I.R = h.rows.R;
            I.ntiles = h.cols.ntiles;
            I.C3 = h.cols3.C;
            I.ntiles3 = h.cols3.ntiles;
            // stage pointers for the k-th step after this build, rotating the
            // three buffers without copies: SRK3 stages (next, cur, base) =
            // (y, x, x), (z, y, x), (y, z, x); Euler (y, x, x); then cur <- y
            for (int k = 0; k < 3; ++k) {
                const int x = (h.cur + k) % 3, y = (h.cur + k + 1) % 3, z = (h.cur + k + 2) % 3;
                const int tab[3][3] = {{y, x, x}, {z, y, x}, {y, z, x}};
                for (int s = 0; s < stages; ++s)
                    ptrs[static_cast<size_t>((k * stages + s) * n + i)] = stage_ptrs<T>(h, tab[s][0], tab[s][1], tab[s][2]);
            }
            const int nb[3] = {h.desc.n_in > 0 ? h.cols.ntiles * h.desc.n_in : 0,
                               (h.desc.n_in > 0 || h.desc.n_out > 0) ? h.rows.nblocks : 0,
                               h.cols3.ntiles * h.desc.ngroups};
            for (int kk = 0; kk < 3; ++kk) {
                firsts[kk][static_cast<size_t>(i)] = static_cast<int>(maps[kk].size());
                for (int b = 0; b < nb[kk]; ++b) maps[kk].push_back(i);
            }
            for (int f : h.static_fields)
                statics.push_back({static_cast<cx<T>*>(h.state[0]) + static_cast<size_t>(f) * h.field_elems,
                                   h.fp32 ? h.k0[0] + f : nullptr, h.desc.nay});
            lc.lds[0] = std::max(lc.lds[0], h.cols.lds);
            lc.lds[1] = std::max(lc.lds[1], h.rows.lds);
            lc.lds[2] = std::max(lc.lds[2], h.cols3.lds3);
        }
        // forward-column workgroup order: momentum groups first when the launch
        // is at most four rounds of the workgroups the GPU holds at once (an
        // occupancy query that fails keeps the mixed order; either order gives
        // the same results)
        {
            const ClassKey& K = lc.key;
            K3Fn<T> k3 = pick_k3<T>(K.c3NT, K.c3V, K.acc, K.c3fz, K.c3nb, K.c3ws);
            int per_cu = 0, ncu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k3), K.c3NT,
                                                             lc.lds[2]) != hipSuccess ||
                hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
                per_cu = ncu = 0;
                (void)hipGetLastError();   // not left for the next launch check to report
            }
            const int nb3 = static_cast<int>(maps[2].size());
            const int order = (per_cu > 0 && ncu > 0 && nb3 <= 4 * per_cu * ncu) ? 1 : 0;
            for (auto& I : inst) I.k3_order = order;
        }
        SHHD_HIP(&err, hipMalloc(&lc.d_inst, inst.size() * sizeof(InstDev<T>)));
        SHHD_HIP(&err, hipMemcpy(lc.d_inst, inst.data(), inst.size() * sizeof(InstDev<T>), hipMemcpyHostToDevice));
        SHHD_HIP(&err, hipMalloc(&lc.d_ptrs, ptrs.size() * sizeof(StagePtrs<T>)));
        SHHD_HIP(&err, hipMemcpy(lc.d_ptrs, ptrs.data(), ptrs.size() * sizeof(StagePtrs<T>), hipMemcpyHostToDevice));
        size_t total = 0;
        for (int k = 0; k < 3; ++k) total += maps[k].size() + firsts[k].size();
        SHHD_HIP(&err, hipMalloc(&lc.d_map, std::max<size_t>(1, total) * sizeof(int)));
        size_t off = 0;
        for (int k = 0; k < 3; ++k) {
            lc.nblk[k] = static_cast<int>(maps[k].size());
            lc.map[k].inst = lc.d_map + off;
            if (!maps[k].empty())
                SHHD_HIP(&err, hipMemcpy(lc.d_map + off, maps[k].data(), maps[k].size() * sizeof(int), hipMemcpyHostToDevice));
            off += maps[k].size();
            lc.map[k].first = lc.d_map + off;
            SHHD_HIP(&err, hipMemcpy(lc.d_map + off, firsts[k].data(), firsts[k].size() * sizeof(int), hipMemcpyHostToDevice));
            off += firsts[k].size();
        }
        // on-device accumulators
        std::vector<MeasDev<T>> meas(static_cast<size_t>(n));
        lc.has_probe = lc.has_sk = false;
        lc.sk_blocks = 0;
        for (int i = 0; i < n; ++i) {
            const int slot = lc.members[static_cast<size_t>(i)];
            Instance& h = *slots[static_cast<size_t>(slot)];
            MeasDev<T>& M = meas[static_cast<size_t>(i)];
            M = MeasDev<T>{};
            M.nprobe = static_cast<int>(h.probes.size());
            M.capacity = h.capacity;
            M.probes = h.d_probes;
            M.ring = h.d_ring;
            M.rows = h.d_rows;
            M.flux = static_cast<const cx<T>*>(h.flux);
            M.sk_nf = static_cast<int>(h.sk_fields.size());
            for (int q = 0; q < M.sk_nf; ++q) M.sk_field[q] = h.sk_fields[static_cast<size_t>(q)];
            M.sk_nevery = std::max(1, h.sk_nevery);
            M.sk_base = h.sk_base;
            M.sk = h.d_sk;
            M.sk_count = h.d_sk_count;
            M.step = d_steps + slot;
            M.flag = d_flags + slot;
            M.field_elems = h.field_elems;
            M.static_mask = 0;
            for (int f : h.static_fields) M.static_mask |= 1u << f;
            M.k0_static = h.fp32 ? h.k0[0] : nullptr;
            lc.has_probe = lc.has_probe || M.nprobe > 0;
            if (M.sk_nf > 0) {
                lc.has_sk = true;
                const size_t total = static_cast<size_t>(M.sk_nf) * h.field_elems;
                lc.sk_blocks = std::max(lc.sk_blocks, static_cast<int>(std::min<size_t>(1024, (total + 255) / 256)));
            }
        }
        if (lc.has_probe || lc.has_sk) {
            SHHD_HIP(&err, hipMalloc(&lc.d_meas, meas.size() * sizeof(MeasDev<T>)));
            SHHD_HIP(&err, hipMemcpy(lc.d_meas, meas.data(), meas.size() * sizeof(MeasDev<T>), hipMemcpyHostToDevice));
        }
        lc.nstatic = static_cast<int>(statics.size());
        if (lc.nstatic) {
            SHHD_HIP(&err, hipMalloc(&lc.d_static, statics.size() * sizeof(StaticEntry<T>)));
            SHHD_HIP(&err, hipMemcpy(lc.d_static, statics.data(), statics.size() * sizeof(StaticEntry<T>), hipMemcpyHostToDevice));
        }
        const ClassKey& K = lc.key;
        SHHD_HIP(&err, hipFuncSetAttribute(reinterpret_cast<const void*>(pick_k1<T>(K.cNT, K.cV, K.cfz, K.cnb)),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lc.lds[0])));
        SHHD_HIP(&err, hipFuncSetAttribute(reinterpret_cast<const void*>(pick_k2<T>(K.rNT, K.rV, K.maxn, K.rfl, K.rnb, K.rlay, K.rws)),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lc.lds[1])));
        SHHD_HIP(&err, hipFuncSetAttribute(reinterpret_cast<const void*>(pick_k3<T>(K.c3NT, K.c3V, K.acc, K.c3fz, K.c3nb, K.c3ws)),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lc.lds[2])));
        return SHHD_OK;
    }

    shhd_status rebuild() {
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        drop_graph();
        rot = 0;
        for (auto& c : classes) c.release();
        classes.clear();
        std::map<ClassKey, std::vector<int>> groups;
        for (int s = 0; s < kMaxSlots; ++s) {
            const Instance* h = slots[static_cast<size_t>(s)].get();
            if (!h || h->failed || h->held) continue;   // a stopped or held instance is in no launch
            ClassKey k{h->fp32,      h->stages,     h->cols.NT,     h->cols.V,       h->cols3.NT, h->cols3.V,
                       h->cols3.acc, h->rows.NT,    h->rows.V,      h->rows.maxn,    h->rows.fl,  h->cols.fz,
                       h->cols.nbuf, h->cols3.fz,   h->cols3.nbuf,  h->rows.nbuf, h->rows.layout, h->cols3.ws,
                       h->rows.ws};
            groups[k].push_back(s);
        }
        std::vector<int> active(kMaxSlots, 0);
        for (auto& g : groups) {
            LaunchClass lc;
            lc.key = g.first;
            lc.members = g.second;
            classes.push_back(lc);
            for (int s : g.second) active[static_cast<size_t>(s)] = 1;
        }
        SHHD_HIP(&err, hipMemcpy(d_active, active.data(), kMaxSlots * sizeof(int), hipMemcpyHostToDevice));
        for (auto& lc : classes) {
            shhd_status st = lc.key.fp32 ? build_class_tables<float>(lc) : build_class_tables<double>(lc);
            if (st != SHHD_OK) return st;
        }
        dirty = false;
        return SHHD_OK;
    }

    hipEvent_t take_event() {
        if (!ev_pool.empty()) {
            hipEvent_t e = ev_pool.back();
            ev_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }

    template <typename F>
    shhd_status timed(int cls, bool capturing, F&& launch) {
        const bool t = timing && !capturing;
        hipEvent_t a = nullptr, b = nullptr;
        if (t) {
            a = take_event();
            b = take_event();
            SHHD_HIP(&err, hipEventRecord(a, ls));
        }
        launch();
        SHHD_HIP(&err, hipGetLastError());
        if (t) {
            SHHD_HIP(&err, hipEventRecord(b, ls));
            ev_used.push_back({cls, {a, b}});
        }
        return SHHD_OK;
    }

    template <typename T>
    shhd_status launch_stage(LaunchClass& lc, int k, int s, const StageParams& sp, bool capturing) {
        const ClassKey& K = lc.key;
        const int n = static_cast<int>(lc.members.size());
        const InstDev<T>* inst = static_cast<const InstDev<T>*>(lc.d_inst);
        const StagePtrs<T>* ptrs = static_cast<const StagePtrs<T>*>(lc.d_ptrs) + static_cast<size_t>(k * K.stages + s) * n;
        shhd_status st = SHHD_OK;
        if (lc.nblk[0] > 0) {
            K1Fn<T> k1 = pick_k1<T>(K.cNT, K.cV, K.cfz, K.cnb);
            st = timed(0, capturing, [&] {
                hipLaunchKernelGGL(k1, dim3(lc.nblk[0]), dim3(K.cNT), lc.lds[0], ls, inst, ptrs, lc.map[0]);
            });
            if (st != SHHD_OK) return st;
        }
        if (lc.nblk[1] > 0) {
            K2Fn<T> k2 = pick_k2<T>(K.rNT, K.rV, K.maxn, K.rfl, K.rnb, K.rlay, K.rws);
            st = timed(1, capturing, [&] {
                hipLaunchKernelGGL(k2, dim3(lc.nblk[1]), dim3(K.rNT), lc.lds[1], ls, inst, lc.map[1]);
            });
            if (st != SHHD_OK) return st;
        }
        K3Fn<T> k3 = pick_k3<T>(K.c3NT, K.c3V, K.acc, K.c3fz, K.c3nb, K.c3ws);
        return timed(2, capturing, [&] {
            hipLaunchKernelGGL(k3, dim3(lc.nblk[2]), dim3(K.c3NT), lc.lds[2], ls, inst, ptrs, sp, lc.map[2]);
        });
    }

    // one step of every instance, k = steps since the tables were built (mod 3)
    shhd_status one_step(int k, bool capturing) {
        static constexpr double S2 = 1.41421356237309504880;
        static constexpr double S3 = 1.73205080756887729352;
        static constexpr double B[3] = {(2.0 * S2 + S3) / 5.0, (-4.0 * S2 + 3.0 * S3) / 5.0, (S2 - 2.0 * S3) / 10.0};
        // time_integrator_srk3_*.h: (w_base, w_current) = (0,1), (3/4,1/4), (1/3,2/3)
        static constexpr double WB[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
        static constexpr double WC[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
        // fork: class c runs on stream c mod nls
        const int nls = lane_streams();
        if (nls > 1) {
            const shhd_status st = ensure_side_streams();
            if (st != SHHD_OK) return st;
            SHHD_HIP(&err, hipEventRecord(fork_ev, stream));
            for (int i = 0; i < nls - 1; ++i) SHHD_HIP(&err, hipStreamWaitEvent(side[i], fork_ev, 0));
        }
        for (size_t c = 0; c < classes.size(); ++c) {
            LaunchClass& lc = classes[c];
            if (!lc.nstatic) continue;
            ls = class_stream(c, nls);
            if (lc.key.fp32)
                hipLaunchKernelGGL(k_enforce_static<float>, dim3(lc.nstatic), dim3(256), 0, ls,
                                   static_cast<const StaticEntry<float>*>(lc.d_static));
            else
                hipLaunchKernelGGL(k_enforce_static<double>, dim3(lc.nstatic), dim3(256), 0, ls,
                                   static_cast<const StaticEntry<double>*>(lc.d_static));
            SHHD_HIP(&err, hipGetLastError());
        }
        for (int s = 0; s < 3; ++s) {
            for (size_t c = 0; c < classes.size(); ++c) {
                LaunchClass& lc = classes[c];
                if (s >= lc.key.stages) continue;
                ls = class_stream(c, nls);
                // stage weights of the step flux: 1/6, 1/6, 2/3 (SRK3) or 1 (Euler)
                static constexpr double FW[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
                StageParams sp{};
                if (lc.key.stages == 3) {
                    sp.beta = B[s];
                    sp.wb = WB[s];
                    sp.wc = WC[s];
                    sp.use_b = 1;
                    sp.flux_w = FW[s];
                } else {
                    sp.beta = 0.0;
                    sp.wb = 0.0;
                    sp.wc = 1.0;
                    sp.use_b = 0;
                    sp.flux_w = 1.0;
                }
                sp.flux_first = s == 0 ? 1 : 0;
                // SRK3 uses draws A and B of the step at every stage; the
                // counter-based generator gives each stage the same values, so
                // every stage draws them (measured faster than keeping them in HBM
                // for stages 2 and 3, in both precisions)
                sp.flux_last = s == lc.key.stages - 1 ? 1 : 0;
                shhd_status st = lc.key.fp32 ? launch_stage<float>(lc, k, s, sp, capturing)
                                             : launch_stage<double>(lc, k, s, sp, capturing);
                if (st != SHHD_OK) return st;
            }
        }
        for (size_t c = 0; c < classes.size(); ++c) {
            LaunchClass& lc = classes[c];
            if (!lc.d_meas) continue;
            ls = class_stream(c, nls);
            shhd_status st = lc.key.fp32 ? launch_measures<float>(lc, k) : launch_measures<double>(lc, k);
            if (st != SHHD_OK) return st;
        }
        // join: the step counters advance once every class has finished the step
        if (nls > 1) {
            for (int i = 0; i < nls - 1; ++i) {
                SHHD_HIP(&err, hipEventRecord(join_ev[i], side[i]));
                SHHD_HIP(&err, hipStreamWaitEvent(stream, join_ev[i], 0));
            }
        }
        ls = stream;
        hipLaunchKernelGGL(k_advance, dim3((kMaxSlots + 255) / 256), dim3(256), 0, stream, d_steps, d_flags, d_active,
                           kMaxSlots);
        SHHD_HIP(&err, hipGetLastError());
        return SHHD_OK;
    }

    // probes and structure factor on the state the step just wrote
    template <typename T>
    shhd_status launch_measures(LaunchClass& lc, int k) {
        const int n = static_cast<int>(lc.members.size());
        const MeasDev<T>* meas = static_cast<const MeasDev<T>*>(lc.d_meas);
        const StagePtrs<T>* last =
            static_cast<const StagePtrs<T>*>(lc.d_ptrs) + static_cast<size_t>(k * lc.key.stages + lc.key.stages - 1) * n;
        if (lc.has_probe) hipLaunchKernelGGL(k_probe<T>, dim3(n), dim3(256), 0, ls, meas, last);
        if (lc.has_sk)
            hipLaunchKernelGGL(k_structure_factor<T>, dim3(lc.sk_blocks, n), dim3(256), 0, ls, meas, last);
        SHHD_HIP(&err, hipGetLastError());
        return SHHD_OK;
    }

    shhd_status step(int nsteps) {
        if (nsteps < 0) return set_err(&err, SHHD_ERR_INVALID, "nsteps must be nonnegative");
        if (nsteps == 0 || active_count() == 0) return SHHD_OK;
        for (auto& sp : slots) {
            if (sp && sp->rho0_needed && sp->desc.rho0 == 0.0)
                return set_err(&err, SHHD_ERR_RUNTIME, "reference density must be nonzero.");
        }
        for (auto& sp : slots) {
            if (sp && sp->capacity && sp->probe_rows + nsteps > sp->capacity)
                return set_err(&err, SHHD_ERR_FULL, "probe ring would overflow: " + std::to_string(sp->probe_rows) +
                                                         " of " + std::to_string(sp->capacity) +
                                                         " rows recorded; read the probes first");
        }
        if (!sticky_failures) {
            for (auto& sp : slots) {
                if (sp && sp->failed) {
                    sp->failed = false;   // a new call retries from the failing step's start
                    dirty = true;
                }
            }
        }
        if (dirty) {
            shhd_status st = rebuild();
            if (st != SHHD_OK) return st;
        }
        if (lane_streams() > 1) {   // created before any graph capture begins
            const shhd_status st = ensure_side_streams();
            if (st != SHHD_OK) return st;
        }
        for (auto& sp : slots) {
            if (sp && sp->flux && sp->flux_clear) {
                SHHD_HIP(&err, hipMemsetAsync(sp->flux, 0, sp->flux_slots() * sp->field_elems * sp->elem, stream));
                sp->flux_clear = false;
            }
        }
        if (!sticky_failures) {
            hipLaunchKernelGGL(k_clear_flags, dim3((kMaxSlots + 255) / 256), dim3(256), 0, stream, d_flags, kMaxSlots);
            SHHD_HIP(&err, hipGetLastError());
        }
        int done = 0;
        auto eager = [&]() {
            const shhd_status st = one_step(rot, false);
            rot = (rot + 1) % 3;
            ++done;
            return st;
        };
        if (use_graph && !timing && nsteps >= 3) {
            if (!graph_exec) {
                SHHD_HIP(&err, hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                shhd_status st = SHHD_OK;
                for (int k = 0; k < 3 && st == SHHD_OK; ++k) st = one_step(k, true);
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(stream, &g);
                if (st != SHHD_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return st;
                }
                if (ec != hipSuccess) return set_err(&err, SHHD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
                graph = g;
                SHHD_HIP(&err, hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
            }
            // the graph holds rotations 0, 1, 2: single steps up to rotation 0 first
            while (rot != 0 && done < nsteps) {
                const shhd_status st = eager();
                if (st != SHHD_OK) return st;
            }
            for (; done + 3 <= nsteps; done += 3) SHHD_HIP(&err, hipGraphLaunch(graph_exec, stream));
        }
        while (done < nsteps) {
            const shhd_status st = eager();
            if (st != SHHD_OK) return st;
        }
        std::vector<int> flags(kMaxSlots);
        std::vector<uint64_t> counts(kMaxSlots);
        SHHD_HIP(&err, hipMemcpyAsync(flags.data(), d_flags, kMaxSlots * sizeof(int), hipMemcpyDeviceToHost, stream));
        SHHD_HIP(&err, hipMemcpyAsync(counts.data(), d_steps, kMaxSlots * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
        SHHD_HIP(&err, hipStreamSynchronize(stream));
        // An instance whose density reached zero completed only the steps its
        // device counter shows; its state is the failing step's start, in the
        // buffer that step read (no stage writes it).
        std::string failed_now;
        for (int s = 0; s < kMaxSlots; ++s) {
            Instance* sp = slots[static_cast<size_t>(s)].get();
            if (!sp) continue;
            uint64_t advanced = sp->held ? 0 : static_cast<uint64_t>(nsteps);
            if (flags[static_cast<size_t>(s)]) {
                advanced = counts[static_cast<size_t>(s)] - sp->step_index;
                if (!sp->failed) {
                    failed_now += (failed_now.empty() ? "" : ", ") + std::string("slot ") + std::to_string(s) +
                                  " at step " + std::to_string(counts[static_cast<size_t>(s)]);
                    sp->failed = true;
                    dirty = true;   // the launch tables drop it
                }
            }
            sp->cur = static_cast<int>((static_cast<uint64_t>(sp->cur) + advanced) % 3);
            sp->step_index += advanced;
            if (sp->capacity) sp->probe_rows += static_cast<int>(advanced);
        }
        if (timing) {
            for (auto& u : ev_used) {
                float ms = 0.f;
                SHHD_HIP(&err, hipEventElapsedTime(&ms, u.second.first, u.second.second));
                time_ms[u.first] += ms;
                launches[u.first] += 1;
                ev_pool.push_back(u.second.first);
                ev_pool.push_back(u.second.second);
            }
            ev_used.clear();
        }
        if (!failed_now.empty()) {
            // the reference's message (src/fcalculator_workspace.cc:126-128); a
            // batch names the stopped instances
            std::string msg = "compressible velocity encountered zero density.";
            if (sticky_failures) msg += " (" + failed_now + ")";
            return set_err(&err, SHHD_ERR_ZERO_DENSITY, msg);
        }
        return SHHD_OK;
    }

    // hold: the instance keeps its state and takes no step until released
    shhd_status hold(int slot, bool on) {
        Instance* h = at(slot);
        if (!h) return set_err(&err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
        if (h->held != on) {
            h->held = on;
            dirty = true;
        }
        return SHHD_OK;
    }

    // instances stopped by zero density: slot and the 0-based step that failed
    shhd_status failures(int* out_slots, long long* out_steps, int capacity, int* count) const {
        int n = 0;
        for (int s = 0; s < kMaxSlots; ++s) {
            const Instance* sp = slots[static_cast<size_t>(s)].get();
            if (!sp || !sp->failed) continue;
            if (n < capacity) {
                if (out_slots) out_slots[n] = s;
                if (out_steps) out_steps[n] = static_cast<long long>(sp->step_index);
            }
            ++n;
        }
        if (count) *count = n;
        return SHHD_OK;
    }
};

// ---------------------------------------------------------------------------
// autotuner: for one problem shape, time every candidate tile of the two
// column passes and the candidate row-pass lane counts on a scratch instance
// (uniform density, the configuration's own noise and terms), keep the fastest
// of each kernel class and record it for the process
// ---------------------------------------------------------------------------
shhd_status time_classes(Engine& e, int steps, double ms[kClasses]) {
    shhd_status st = e.step(2);   // builds the tables and graphs of the shape
    if (st != SHHD_OK) return st;
    e.timing = true;
    for (int k = 0; k < kClasses; ++k) {
        e.time_ms[k] = 0.0;
        e.launches[k] = 0;
    }
    st = e.step(steps);
    e.timing = false;
    for (int k = 0; k < kClasses; ++k) ms[k] = e.launches[k] ? e.time_ms[k] / static_cast<double>(e.launches[k]) : 0.0;
    return st;
}

shhd_status autotune_config(const shhd_config& cfg, int steps, Tuning& best, double best_ms[kClasses],
                            std::string* err) {
    Engine e;
    shhd_status st = e.init(0, false, false);
    if (st != SHHD_OK) return set_err(err, st, e.err);
    int slot = -1;
    st = e.admit(cfg, nullptr, 0, &slot);
    if (st != SHHD_OK) return set_err(err, st, e.err);
    shhd_initial_condition rho{};
    rho.kind = SHHD_IC_EQUILIBRIUM_DENSITY;
    rho.value = 1.0;
    rho.pressure_coefficient = 1.0;
    st = e.apply_initial_condition(slot, &rho);
    if (st != SHHD_OK) return set_err(err, st, e.err);
    Instance& h = *e.at(slot);
    const Desc& d = h.desc;

    // distinct candidate plans of each pass
    auto col_candidates = [&](int seg) {
        std::vector<int> out;
        std::vector<int> seen;
        for (int c : {1, 2, 4, 8}) {
            const ColPlan p = plan_columns(d.my, d.nay, d.hxa, h.elem, 0, seg, c);
            if (std::find(seen.begin(), seen.end(), p.C) != seen.end()) continue;
            seen.push_back(p.C);
            out.push_back(p.C);
        }
        return out;
    };
    std::vector<int> row_nts;
    {
        const RowPlan p0 = plan_rows(d.mx, d.my, h.elem, 0);
        row_nts.push_back(p0.NT);
        if (p0.NT < 1024) {
            const RowPlan p1 = plan_rows(d.mx, d.my, h.elem, p0.NT == 64 ? 256 : 2 * p0.NT);
            if (p1.NT != p0.NT && p1.R > p0.R) row_nts.push_back(p1.NT);
        }
    }
    Tuning cur = default_tuning();
    double ms[kClasses];
    // inverse column tile
    double bk = 1e30;
    for (int c : col_candidates(32)) {
        Tuning t = cur;
        t.c1 = c;
        if ((st = e.replan(slot, t)) != SHHD_OK || (st = time_classes(e, steps, ms)) != SHHD_OK) return set_err(err, st, e.err);
        if (d.n_in > 0 && ms[0] < bk) { bk = ms[0]; best.c1 = c; best_ms[0] = ms[0]; }
    }
    cur.c1 = best.c1;
    // forward column tile
    bk = 1e30;
    for (int c : col_candidates(16)) {
        Tuning t = cur;
        t.c3 = c;
        if ((st = e.replan(slot, t)) != SHHD_OK || (st = time_classes(e, steps, ms)) != SHHD_OK) return set_err(err, st, e.err);
        if (ms[2] < bk) { bk = ms[2]; best.c3 = c; best_ms[2] = ms[2]; }
    }
    cur.c3 = best.c3;
    // column transforms (shared by both column passes: their summed time)
    bk = 1e30;
    for (int ct = 0; ct < kTransforms; ++ct) {
        Tuning t = cur;
        t.col_transform = ct;
        if ((st = e.replan(slot, t)) != SHHD_OK || (st = time_classes(e, steps, ms)) != SHHD_OK) return set_err(err, st, e.err);
        if (ms[0] + ms[2] < bk) { bk = ms[0] + ms[2]; best.col_transform = ct; best_ms[0] = ms[0]; best_ms[2] = ms[2]; }
    }
    cur.col_transform = best.col_transform;
    // row-pass lanes, then row transforms
    bk = 1e30;
    for (int nt : row_nts) {
        Tuning t = cur;
        t.rnt = nt;
        if ((st = e.replan(slot, t)) != SHHD_OK || (st = time_classes(e, steps, ms)) != SHHD_OK) return set_err(err, st, e.err);
        if (ms[1] < bk) { bk = ms[1]; best.rnt = nt; best_ms[1] = ms[1]; }
    }
    cur.rnt = best.rnt;
    bk = 1e30;
    for (int rt = 0; rt < kTransforms; ++rt) {
        Tuning t = cur;
        t.row_transform = rt;
        if ((st = e.replan(slot, t)) != SHHD_OK || (st = time_classes(e, steps, ms)) != SHHD_OK) return set_err(err, st, e.err);
        if (ms[1] < bk) { bk = ms[1]; best.row_transform = rt; best_ms[1] = ms[1]; }
    }
    cur.row_transform = best.row_transform;
    // kx blocks of the row-to-column planes: the row pass writes them, the
    // forward column pass reads them (their summed time)
    bk = 1e30;
    for (int ulg = 0; ulg < 4; ++ulg) {
        if (d.hxa % (1 << ulg) != 0) break;
        Tuning t = cur;
        t.ulg = ulg;
        if ((st = e.replan(slot, t)) != SHHD_OK || (st = time_classes(e, steps, ms)) != SHHD_OK) return set_err(err, st, e.err);
        if (ms[1] + ms[2] < bk) { bk = ms[1] + ms[2]; best.ulg = ulg; best_ms[1] = ms[1]; best_ms[2] = ms[2]; }
    }
    cur.ulg = best.ulg;
    // kx blocks of the column-to-row planes: the inverse column pass writes
    // them, the row pass reads them (their summed time)
    bk = 1e30;
    for (int tlg = -1; tlg < 4; ++tlg) {
        if (tlg >= 0 && d.hxa % (1 << tlg) != 0) break;
        Tuning t = cur;
        t.tlg = tlg;
        if ((st = e.replan(slot, t)) != SHHD_OK || (st = time_classes(e, steps, ms)) != SHHD_OK) return set_err(err, st, e.err);
        if (ms[0] + ms[1] < bk) { bk = ms[0] + ms[1]; best.tlg = tlg; best_ms[0] = ms[0]; best_ms[1] = ms[1]; }
    }
    cur.tlg = best.tlg;
    // forward-column wave groups (one-wave fused tiles only; elsewhere both
    // choices plan the same launch)
    bk = 1e30;
    for (int ws : {1, 4}) {
        Tuning t = cur;
        t.k3ws = ws;
        if ((st = e.replan(slot, t)) != SHHD_OK || (st = time_classes(e, steps, ms)) != SHHD_OK) return set_err(err, st, e.err);
        if (ms[2] < bk) { bk = ms[2]; best.k3ws = ws; best_ms[2] = ms[2]; }
    }
    cur.k3ws = best.k3ws;
    // row-pass wave groups (one-wave staged row tiles only)
    bk = 1e30;
    for (int ws : {1, 4}) {
        Tuning t = cur;
        t.k2ws = ws;
        if ((st = e.replan(slot, t)) != SHHD_OK || (st = time_classes(e, steps, ms)) != SHHD_OK) return set_err(err, st, e.err);
        if (ms[1] < bk) { bk = ms[1]; best.k2ws = ws; best_ms[1] = ms[1]; }
    }
    store_tuning(h.tune_key, best);
    return SHHD_OK;
}

double kernel_bytes_of(const Instance& h, int cls) {
    const Desc& d = h.desc;
    const double fb = static_cast<double>(h.field_elems) * static_cast<double>(h.elem);
    const double pb = static_cast<double>(h.plane_elems) * static_cast<double>(h.elem);
    if (cls == 0) return d.n_in * (fb + pb);
    if (cls == 1) return (d.n_in + d.n_out) * pb;
    // the base state is read only by the stages that weight it (SRK3 stages
    // 2 and 3; never in Euler): its share of a launch, averaged over a step
    const double base = h.stages == 3 ? 2.0 / 3.0 : 0.0;
    double k3 = d.n_out * pb;
    for (int g = 0; g < d.ngroups; ++g) {
        const Group& G = d.groups[g];
        if (G.kind == GROUP_PSI) {
            k3 += (2.0 + base) * fb;   // cur, next (and base) of psi_a
            k3 += static_cast<double>(d.n_igrad > 0 ? 2 : 0) * fb;
        } else if (d.mode == SHHD_MODE_COMPRESSIBLE) {
            k3 += 3.0 * (2.0 + base) * fb;   // rho, jx, jy: cur, next (and base)
        } else {
            k3 += 2.0 * (2.0 + base) * fb + (d.pressure_coefficient != 0.0 ? fb : 0.0);
        }
    }
    return k3;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI: single simulation
// ---------------------------------------------------------------------------
struct shhd_gpu {
    Engine eng;
    int slot = -1;
    bool state_set = false;
    Instance* inst() { return eng.at(slot); }
};

struct shhd_batch {
    Engine eng;
};

extern "C" {

shhd_status shhd_gpu_create(const shhd_config* config, shhd_gpu** out) {
    if (!config || !out) return set_err(nullptr, SHHD_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<Instance> probe;
    shhd_status st = prepare(*config, false, probe);   // configuration errors before any device work
    if (st != SHHD_OK) return st;
    auto h = std::make_unique<shhd_gpu>();
    // steps of a call are replayed from a captured 3-step hipGraph
    // (SHHD_USE_GRAPH=0 launches every kernel eagerly instead)
    const char* g = std::getenv("SHHD_USE_GRAPH");
    st = h->eng.init(0, false, !g || std::atoi(g) != 0);
    if (st != SHHD_OK) return set_err(nullptr, st, h->eng.err);
    st = h->eng.admit(*config, nullptr, 0, &h->slot);
    if (st != SHHD_OK) return set_err(nullptr, st, h->eng.err);
    if (std::getenv("SHHD_VERBOSE")) {
        const Instance& I = *h->inst();
        const Desc& d = I.desc;
        std::fprintf(stderr,
                     "[shhd] Mx=%d My=%d Na=%dx%d N=%d fp%d F_in=%d F_out=%d groups=%d | cols C=%d NT=%d V=%d "
                     "tiles=%d lds=%zu | cols3 C=%d NT=%d acc=%d lds=%zu | rows R=%d NT=%d V=%d blocks=%d maxn=%d lds=%zu\n",
                     d.mx, d.my, d.nax, d.nay, d.N, I.fp32 ? 32 : 64, d.n_in, d.n_out, d.ngroups, I.cols.C, I.cols.NT,
                     I.cols.V, I.cols.ntiles, I.cols.lds, I.cols3.C, I.cols3.NT, I.cols3.acc, I.cols3.lds3, I.rows.R,
                     I.rows.NT, I.rows.V, I.rows.nblocks, I.rows.maxn, I.rows.lds);
    }
    *out = h.release();
    return SHHD_OK;
}

shhd_status shhd_autotune(const shhd_config* config, int steps, shhd_tuning* out) {
    if (!config || !out || steps <= 0) return set_err(nullptr, SHHD_ERR_INVALID, "invalid autotune arguments");
    Tuning best{};
    double ms[kClasses] = {0, 0, 0};
    std::string err;
    const shhd_status st = autotune_config(*config, steps, best, ms, &err);
    if (st != SHHD_OK) return set_err(nullptr, st, err);
    out->inverse_column_tile = best.c1;
    out->forward_column_tile = best.c3;
    out->row_lanes = best.rnt;
    out->column_transform = best.col_transform;
    out->row_transform = best.row_transform;
    out->plane_block = best.ulg;
    out->inverse_plane_block = best.tlg + 1;
    out->forward_column_waves = best.k3ws;
    out->row_waves = best.k2ws;
    for (int k = 0; k < kClasses; ++k) out->ms[k] = ms[k];
    return SHHD_OK;
}

shhd_status shhd_set_tuning(const shhd_config* config, const shhd_tuning* t) {
    if (!config || !t) return set_err(nullptr, SHHD_ERR_INVALID, "null argument");
    std::unique_ptr<Instance> h;
    const Tuning none{};
    shhd_status st = prepare(*config, false, h, &none);
    if (st != SHHD_OK) return st;
    if (t->inverse_column_tile < 0 || t->forward_column_tile < 0 ||
        (t->row_lanes != 0 && t->row_lanes != 64 && t->row_lanes != 256 && t->row_lanes != 512 && t->row_lanes != 1024) ||
        t->column_transform < 0 || t->column_transform >= kTransforms || t->row_transform < 0 ||
        t->row_transform >= kTransforms || t->plane_block < 0 || t->plane_block > 12 || t->inverse_plane_block < 0 ||
        t->inverse_plane_block > 13 ||
        (t->forward_column_waves != 0 && t->forward_column_waves != 1 && t->forward_column_waves != 4) ||
        (t->row_waves != 0 && t->row_waves != 1 && t->row_waves != 4))
        return set_err(nullptr, SHHD_ERR_INVALID, "invalid launch shapes");
    store_tuning(h->tune_key, Tuning{t->inverse_column_tile, t->forward_column_tile, t->row_lanes, t->column_transform,
                                     t->row_transform, t->plane_block, t->inverse_plane_block - 1,
                                     t->forward_column_waves, t->row_waves});
    return SHHD_OK;
}

shhd_status shhd_gpu_set_state(shhd_gpu* h, const double* in, size_t n) {
    if (!h || !in) return set_err(h ? &h->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    Instance* I = h->inst();
    shhd_status st = upload_state(I, in, n, h->eng.stream, &h->eng.err);
    if (st != SHHD_OK) return st;
    h->state_set = true;
    h->eng.dirty = true;   // rho0 of the descriptor may have changed
    return SHHD_OK;
}

shhd_status shhd_gpu_get_state(shhd_gpu* h, double* out, size_t n) {
    if (!h || !out) return set_err(h ? &h->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    return download_state(h->inst(), out, n, h->eng.stream, &h->eng.err);
}

shhd_status shhd_gpu_step(shhd_gpu* h, int nsteps, double t0) {
    (void)t0;   // stage times are cache keys only in the reference
    if (!h) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    if (!h->state_set) return set_err(&h->eng.err, SHHD_ERR_INVALID, "shhd_gpu_set_state must be called before stepping");
    return h->eng.step(nsteps);
}

shhd_status shhd_gpu_set_flux_request(shhd_gpu* h, int density, int order_parameter, int momentum) {
    if (!h) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return h->eng.set_flux_request(h->slot, density, order_parameter, momentum);
}

shhd_status shhd_gpu_get_flux(shhd_gpu* h, double* out, size_t n) {
    if (!h || !out) return set_err(h ? &h->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    return h->eng.get_flux(h->slot, out, n);
}

shhd_status shhd_gpu_set_probes(shhd_gpu* h, const shhd_probe* probes, int nprobes, int capacity) {
    if (!h) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return h->eng.set_probes(h->slot, probes, nprobes, capacity);
}

shhd_status shhd_gpu_read_probes(shhd_gpu* h, double* out, size_t n, int* nrows) {
    if (!h) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return h->eng.read_probes(h->slot, out, n, nrows);
}

shhd_status shhd_gpu_set_structure_factor(shhd_gpu* h, unsigned field_mask, int nevery) {
    if (!h) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return h->eng.set_structure_factor(h->slot, field_mask, nevery);
}

shhd_status shhd_gpu_get_structure_factor(shhd_gpu* h, double* out, size_t n, long long* nsamples) {
    if (!h || !out) return set_err(h ? &h->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    return h->eng.get_structure_factor(h->slot, out, n, nsamples);
}

shhd_status shhd_gpu_apply_initial_condition(shhd_gpu* h, const shhd_initial_condition* ic) {
    if (!h) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    const shhd_status st = h->eng.apply_initial_condition(h->slot, ic);
    if (st == SHHD_OK) h->state_set = true;   // the state starts at zero, as a new reference State does
    return st;
}

// SnapshotMeasure::write_physical_snapshot / write_spectral_snapshot
// (src/measure_snapshot.cc:54-138): the same header lines and columns,
// numbers as std::scientific with 16 digits (printf "%.16e").
shhd_status shhd_write_snapshot(const char* path, int space, const double* data, int nx, int ny,
                                int num_order_parameters, int step, double time) {
    if (!path || !data || nx <= 0 || ny <= 0 || num_order_parameters < 0 ||
        (space != SHHD_SNAPSHOT_PHYSICAL && space != SHHD_SNAPSHOT_SPECTRAL))
        return set_err(nullptr, SHHD_ERR_INVALID, "invalid snapshot arguments");
    FILE* fp = std::fopen(path, "w");
    if (!fp) return set_err(nullptr, SHHD_ERR_RUNTIME, std::string("SnapshotMeasure: cannot open file: ") + path);
    const int N = num_order_parameters;
    const bool phys = space == SHHD_SNAPSHOT_PHYSICAL;
    const int nxs = phys ? nx : nx / 2 + 1;
    const size_t field = static_cast<size_t>(nxs) * ny;
    std::string line;
    char num[40];
    std::fprintf(fp, "# snapshot space %s\n", phys ? "physical" : "spectral");
    std::fprintf(fp, "# step %d time %.16e\n", step, time);
    std::fprintf(fp, phys ? "# nx %d ny %d order_parameters %d\n" : "# nkx %d nky %d order_parameters %d\n", nxs, ny, N);
    line = phys ? "# x y rho" : "# kx_index ky_index rho_real rho_imag";
    for (int a = 0; a < N; ++a) {
        line += phys ? " psi_com" + std::to_string(a)
                     : " psi_com" + std::to_string(a) + "_real psi_com" + std::to_string(a) + "_imag";
    }
    line += phys ? " jx jy\n" : " jx_real jx_imag jy_real jy_imag\n";
    std::fputs(line.c_str(), fp);
    const int nf = N + 3;
    std::string buf;
    buf.reserve(1 << 20);
    for (int gy = 0; gy < ny; ++gy) {
        for (int gx = 0; gx < nxs; ++gx) {
            const size_t i = static_cast<size_t>(gy) * nxs + gx;
            buf += std::to_string(gx);
            buf += ' ';
            buf += std::to_string(gy);
            for (int f = 0; f < nf; ++f) {
                if (phys) {
                    std::snprintf(num, sizeof(num), " %.16e", data[static_cast<size_t>(f) * field + i]);
                    buf += num;
                } else {
                    const size_t o = 2 * (static_cast<size_t>(f) * field + i);
                    std::snprintf(num, sizeof(num), " %.16e", data[o]);
                    buf += num;
                    std::snprintf(num, sizeof(num), " %.16e", data[o + 1]);
                    buf += num;
                }
            }
            buf += '\n';
            if (buf.size() > (1u << 20) - 1024) {
                std::fwrite(buf.data(), 1, buf.size(), fp);
                buf.clear();
            }
        }
    }
    std::fwrite(buf.data(), 1, buf.size(), fp);
    const bool ok = std::ferror(fp) == 0;
    if (std::fclose(fp) != 0 || !ok)
        return set_err(nullptr, SHHD_ERR_RUNTIME, std::string("SnapshotMeasure: cannot write file: ") + path);
    return SHHD_OK;
}

shhd_status shhd_gpu_get_physical(shhd_gpu* h, double* out, size_t n) {
    if (!h || !out) return set_err(h ? &h->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    return h->eng.get_physical(h->slot, out, n);
}

void shhd_gpu_destroy(shhd_gpu* h) { delete h; }

const char* shhd_gpu_last_error(const shhd_gpu* h) { return h ? h->eng.err.c_str() : g_create_error.c_str(); }

shhd_status shhd_gpu_get_info(const shhd_gpu* h, shhd_info* info) {
    if (!h || !info) return SHHD_ERR_INVALID;
    const Instance* I = const_cast<shhd_gpu*>(h)->inst();
    info->num_fields = I->desc.nfields;
    info->num_inverse_fields = I->n_in;
    info->num_forward_fields = I->n_out;
    info->num_updated_fields = I->n_updated;
    info->stages_per_step = I->stages;
    info->state_doubles = 2 * ref_spec(I->cfg) * static_cast<size_t>(I->desc.nfields);
    info->device_bytes = I->bytes;
    info->inverse_column_tile = I->cols.C;
    info->forward_column_tile = I->cols3.C;
    info->row_lanes = I->rows.NT;
    info->inverse_column_fused = I->cols.fz > 0 ? I->cols.nbuf : 0;
    info->forward_column_fused = I->cols3.fz > 0 ? I->cols3.nbuf : 0;
    info->row_fused = I->rows.nbuf;
    info->inverse_column_lanes = I->cols.NT;
    info->forward_column_lanes = I->cols3.NT;
    return SHHD_OK;
}

shhd_status shhd_gpu_set_kernel_timing(shhd_gpu* h, int enable) {
    if (!h) return SHHD_ERR_INVALID;
    h->eng.timing = enable != 0;
    for (int i = 0; i < kClasses; ++i) {
        h->eng.time_ms[i] = 0.0;
        h->eng.launches[i] = 0;
    }
    return SHHD_OK;
}

shhd_status shhd_gpu_get_kernel_timing(shhd_gpu* h, double* total_ms, long long* launches, int nclasses) {
    if (!h || !total_ms || !launches) return SHHD_ERR_INVALID;
    for (int i = 0; i < nclasses && i < kClasses; ++i) {
        total_ms[i] = h->eng.time_ms[i];
        launches[i] = h->eng.launches[i];
    }
    return SHHD_OK;
}

shhd_status shhd_gpu_get_kernel_bytes(const shhd_gpu* h, double* bytes, int nclasses) {
    if (!h || !bytes) return SHHD_ERR_INVALID;
    const Instance* I = const_cast<shhd_gpu*>(h)->inst();
    for (int i = 0; i < nclasses && i < kClasses; ++i) bytes[i] = kernel_bytes_of(*I, i);
    return SHHD_OK;
}

// ---------------------------------------------------------------------------
// C-ABI: batched instances
// ---------------------------------------------------------------------------
shhd_status shhd_batch_create(size_t device_budget_bytes, int use_graphs, shhd_batch** out) {
    if (!out) return set_err(nullptr, SHHD_ERR_INVALID, "null argument");
    auto b = std::make_unique<shhd_batch>();
    shhd_status st = b->eng.init(device_budget_bytes, true, use_graphs != 0, true);
    if (st != SHHD_OK) return set_err(nullptr, st, b->eng.err);
    *out = b.release();
    return SHHD_OK;
}

shhd_status shhd_batch_footprint(const shhd_config* config, size_t* bytes) {
    if (!config || !bytes) return set_err(nullptr, SHHD_ERR_INVALID, "null argument");
    std::unique_ptr<Instance> h;
    shhd_status st = prepare(*config, true, h);
    if (st != SHHD_OK) return st;
    *bytes = h->bytes;
    return SHHD_OK;
}

shhd_status shhd_batch_admit(shhd_batch* b, const shhd_config* config, const double* state, size_t n, int* slot) {
    if (!b || !config || !slot) return set_err(b ? &b->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    return b->eng.admit(*config, state, n, slot);   // state NULL: all fields zero
}

shhd_status shhd_batch_step(shhd_batch* b, int nsteps) {
    if (!b) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return b->eng.step(nsteps);
}

shhd_status shhd_batch_hold(shhd_batch* b, int slot, int hold) {
    if (!b) return set_err(nullptr, SHHD_ERR_INVALID, "null batch");
    return b->eng.hold(slot, hold != 0);
}

shhd_status shhd_batch_failures(const shhd_batch* b, int* slots, long long* steps, int capacity, int* count) {
    if (!b || capacity < 0 || !count) return set_err(nullptr, SHHD_ERR_INVALID, "invalid arguments");
    return b->eng.failures(slots, steps, capacity, count);
}

shhd_status shhd_batch_get_state(shhd_batch* b, int slot, double* out, size_t n) {
    if (!b || !out) return set_err(b ? &b->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    const Instance* I = b->eng.at(slot);
    if (!I) return set_err(&b->eng.err, SHHD_ERR_INVALID, "no instance in slot " + std::to_string(slot));
    return download_state(I, out, n, b->eng.stream, &b->eng.err);
}

shhd_status shhd_batch_retire(shhd_batch* b, int slot) {
    if (!b) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return b->eng.retire(slot);
}

shhd_status shhd_batch_set_flux_request(shhd_batch* b, int slot, int density, int order_parameter, int momentum) {
    if (!b) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return b->eng.set_flux_request(slot, density, order_parameter, momentum);
}

shhd_status shhd_batch_get_flux(shhd_batch* b, int slot, double* out, size_t n) {
    if (!b || !out) return set_err(b ? &b->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    return b->eng.get_flux(slot, out, n);
}

shhd_status shhd_batch_set_probes(shhd_batch* b, int slot, const shhd_probe* probes, int nprobes, int capacity) {
    if (!b) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return b->eng.set_probes(slot, probes, nprobes, capacity);
}

shhd_status shhd_batch_read_probes(shhd_batch* b, int slot, double* out, size_t n, int* nrows) {
    if (!b) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return b->eng.read_probes(slot, out, n, nrows);
}

shhd_status shhd_batch_set_structure_factor(shhd_batch* b, int slot, unsigned field_mask, int nevery) {
    if (!b) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return b->eng.set_structure_factor(slot, field_mask, nevery);
}

shhd_status shhd_batch_get_structure_factor(shhd_batch* b, int slot, double* out, size_t n, long long* nsamples) {
    if (!b || !out) return set_err(b ? &b->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    return b->eng.get_structure_factor(slot, out, n, nsamples);
}

shhd_status shhd_batch_apply_initial_condition(shhd_batch* b, int slot, const shhd_initial_condition* ic) {
    if (!b) return set_err(nullptr, SHHD_ERR_INVALID, "null handle");
    return b->eng.apply_initial_condition(slot, ic);
}

shhd_status shhd_batch_apply_initial_condition_noalloc(shhd_batch* b, int slot, const shhd_initial_condition* ic) {
    if (!b) return set_err(nullptr, SHHD_ERR_INVALID, "null batch");
    return b->eng.apply_initial_condition(slot, ic, true);
}

shhd_status shhd_batch_get_physical(shhd_batch* b, int slot, double* out, size_t n) {
    if (!b || !out) return set_err(b ? &b->eng.err : nullptr, SHHD_ERR_INVALID, "null argument");
    return b->eng.get_physical(slot, out, n);
}

shhd_status shhd_batch_usage(const shhd_batch* b, int* active, size_t* used_bytes) {
    if (!b) return SHHD_ERR_INVALID;
    if (active) *active = b->eng.active_count();
    if (used_bytes) *used_bytes = b->eng.used;
    return SHHD_OK;
}

void shhd_batch_destroy(shhd_batch* b) { delete b; }

const char* shhd_batch_last_error(const shhd_batch* b) { return b ? b->eng.err.c_str() : g_create_error.c_str(); }

}  // extern "C"
