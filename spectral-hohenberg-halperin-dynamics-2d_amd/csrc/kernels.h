// kernels.h — the three kernels of one right-hand-side evaluation + stage
// update, templated on the arithmetic type T (float | double), the workgroup
// size NT and the values held per thread VPT.
//
//   K1 k_inverse_columns : compact spectra -> zero-padded inverse y transform
//                          of every needed field; 1/(Mx*My) folded in.
//                          (FCalculatorWorkspace::ensure_physical_fields,
//                          src/fcalculator_workspace.cc:68-145, first half of
//                          FourierTransform2D::backward_many)
//   K2 k_rows            : per row y: inverse x transform (two real fields per
//                          complex FFT), v = j/rho, the pointwise products
//                          (mu_phys, psi v, j (x) v, v, force sources), then
//                          forward x transform truncated to active kx.
//                          (src/fcalculator.cc:189-274, 343-494, 776-840 and
//                          the x halves of forward_many/backward_many)
//   K3 k_forward_columns : forward y transform of the products, truncated to
//                          active ky; spectral RHS (linear mu, pressure,
//                          viscous, advection divergence, forces), in-kernel
//                          noise, stage combination, mask, Hermitian kx = 0
//                          line.  (src/fcalculator.cc:60-187, 496-700,
//                          src/time_integrator.h:70-109 and the
//                          calculate_stage loops of time_integrator_*.h)
#pragma once

#include "../../include/shhd_gpu.h"
#include "desc.h"
#include "fft_fused.h"

#include <type_traits>

namespace shhd {

// Column transforms of K1 and K3: the fixed-length build when the launch
// tuning chose it (and one exists for this length and shape), else the
// runtime plan with the tuned radix order.
template <typename T, int NT, int VPT>
__device__ __forceinline__ void col_fft(cx<T>* buf, int L, int nseq, const Desc& d, const cx<T>* __restrict__ tw,
                                        int sign) {
    if (d.fixed_y) lds_fft_any<T, NT, VPT>(buf, L, nseq, d.plan_y, tw, sign);
    else lds_fft<T, NT, VPT>(buf, L, nseq, d.plan_y, tw, sign);
}

// XCD-aware workgroup order.  The dispatcher hands workgroup b to XCD b % 8,
// and each XCD has its own L2; adjacent column tiles read and write
// neighbouring segments of the same cache lines of the row-major T and U
// planes.  Renumbering so that every XCD gets one contiguous range of tiles
// lets its L2 serve those shared lines once instead of once per tile.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    constexpr int X = 8;
    const int x = b % X, i = b / X, q = nb / X, r = nb % X;
    return x * q + (x < r ? x : r) + i;
}

// Slot i of the n slots [start, start + n) of the forward-column pass, ng
// groups per tile (slot l: tile l / ng, group l mod ng; the momentum group
// last), taken momentum groups first: the slots l = ng - 1 (mod ng), then
// the others in order.  A bijection of the range onto itself.
__device__ __forceinline__ int heavy_first(int start, int n, int i, int ng) {
    const int m0 = start + (ng - 1 - start % ng);   // first momentum slot >= start
    const int nm = m0 < start + n ? (start + n - 1 - m0) / ng + 1 : 0;
    if (i < nm) return m0 + i * ng;
    const int t = (start / ng) * (ng - 1) + min(start % ng, ng - 1) + (i - nm);   // rank among the psi slots
    return (t / (ng - 1)) * ng + t % (ng - 1);
}

// element (kx, y) of a T plane in the blocked layout (Desc::tlg >= 0)
__device__ __forceinline__ uint32_t t_at(const Desc& d, int kx, int y) {
    return ((static_cast<uint32_t>(kx >> d.tlg) * d.my + y) << d.tlg) + (kx & ((1 << d.tlg) - 1));
}

// v / n for n > 0: a shift when ln = log2(n) >= 0 (n a power of two), else a division
__device__ __forceinline__ int divn(int v, int n, int ln) { return ln >= 0 ? (v >> ln) : v / n; }
// log2 of a power of two
__device__ __forceinline__ int log2p2(int n) { return 31 - __clz(n); }

// signed ky index of padded row gy (src/domain.cc:145-149)
__device__ __forceinline__ int signed_gy(int gy, int my) { return (gy <= my / 2) ? gy : gy - my; }

// ---------------------------------------------------------------------------
// K1
// ---------------------------------------------------------------------------
// FZL > 0: a build for columns of exactly FZL points, C = NT * VPT / FZL per
// workgroup, whose first transform pass loads the compact spectra straight
// into its butterflies (no LDS staging) and whose passes alternate between
// NBUF LDS buffers.
template <typename T, int NT, int VPT, int FZL, int NBUF>
__global__ __launch_bounds__(NT) void k_inverse_columns(const InstDev<T>* __restrict__ insts,
                                                        const StagePtrs<T>* __restrict__ ptrs, BlockMap map) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int inst = map.inst[blk];
    const int local = blk - map.first[inst];
    const InstDev<T>& I = insts[inst];
    if (instance_failed(I.flag)) return;
    const Desc& d = I.d;
    const StagePtrs<T>& P = ptrs[inst];
    const int C = I.C;
    const cx<T>* __restrict__ twy = I.twy;
    cx<T>* __restrict__ tbuf = I.tbuf;
    const int slot = local / I.ntiles;
    const cx<T>* __restrict__ src = P.cur[d.in_field[slot]];
    const int gx0 = (local - slot * I.ntiles) * C;
    const int L = d.my;
    const int ncol = C * L;
    const int half = d.nay / 2;
    const T scale = T(d.in_scale);
    const cx<T>* res = lds;   // the transformed columns, natural order
    if constexpr (FZL > 0) {
        constexpr int CF = NT * VPT / FZL;
        constexpr int R0 = fz_radix(FZL, false, 0);
        using M0 = FzMap<NT, FZL, CF, R0>;
        const int tid = threadIdx.x;
        FzBufs<T, NBUF> bf{lds, lds + lp_size(CF * FZL), 0};
        cx<T> v[M0::B][R0];
        // every load is issued before any is used, at an index inside the
        // spectra (the last column past the last one, ky = 0 in the padding),
        // and the padding is zeroed afterwards by selects: a load under a
        // per-point branch is waited for inside its branch, which serializes
        // the column's loads
#pragma unroll
        for (int b = 0; b < M0::B; ++b) {
            const int id = tid + NT * b;
            if (!M0::valid(id)) continue;
            const int gx = gx0 + M0::seq(id);
            const int j = M0::bfly(id);
            const uint32_t col = static_cast<uint32_t>((gx < d.hxa ? gx : d.hxa - 1) * d.nay);
#pragma unroll
            for (int q = 0; q < R0; ++q) {
                const int ky = signed_gy(j + q * M0::M, FZL);
                const bool live = ky < half && -ky < half;
                const int cidx = live ? (ky >= 0 ? ky : ky + d.nay) : 0;
                v[b][q] = gload(src, col + static_cast<uint32_t>(cidx));
            }
        }
#pragma unroll
        for (int b = 0; b < M0::B; ++b) {
            const int id = tid + NT * b;
            if (!M0::valid(id)) continue;
            const bool col_ok = gx0 + M0::seq(id) < d.hxa;
            const int j = M0::bfly(id);
#pragma unroll
            for (int q = 0; q < R0; ++q) {
                const int ky = signed_gy(j + q * M0::M, FZL);
                const bool live = col_ok && ky < half && -ky < half;
                v[b][q] = live ? scale * v[b][q] : czero<T>();
            }
        }
        fz_first<T, NT, FZL, CF, false, NBUF>(bf, v, +1, tid);
        fz_rest_to_lds<T, NT, FZL, CF, false, NBUF>(bf, twy, +1, tid);
        __syncthreads();
        res = bf.rd();
    } else {
        // element e = c L + y of the tile (C L <= NT VPT, host-checked); all
        // loads issued first at indices inside the spectra, masks applied after
        cx<T> v[VPT];
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            const int e = static_cast<int>(threadIdx.x) + NT * i;
            const int c = e / L;
            const int ky = signed_gy(e - c * L, L);
            const int gx = gx0 + c;
            const int cidx = (ky < half && -ky < half) ? (ky >= 0 ? ky : ky + d.nay) : 0;
            v[i] = gload(src, static_cast<uint32_t>((gx < d.hxa ? gx : d.hxa - 1) * d.nay + cidx));
        }
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            const int e = static_cast<int>(threadIdx.x) + NT * i;
            if (e >= ncol) continue;
            const int c = e / L;
            const int ky = signed_gy(e - c * L, L);
            const bool live = gx0 + c < d.hxa && ky < half && -ky < half;
            lds[lp(e)] = live ? scale * v[i] : czero<T>();
        }
        __syncthreads();
        col_fft<T, NT, VPT>(lds, L, C, d, twy, +1);
    }
    cx<T>* __restrict__ dst = tbuf + static_cast<size_t>(slot) * L * d.hxa;
    const int lc = log2p2(C);
    if (d.tlg < 0) {
        // row-major T: the lanes of a row write its C columns side by side
        for (int idx = threadIdx.x; idx < ncol; idx += NT) {
            const int y = idx >> lc;
            const int c = idx - (y << lc);
            const int gx = gx0 + c;
            if (gx < d.hxa) gstore(dst, static_cast<uint32_t>(y * d.hxa + gx), res[lp(c * L + y)]);
        }
    } else {
        // kx-blocked T: the tile's columns of one block, all rows, are one
        // contiguous run of the plane (lb = log2 of the columns a run holds)
        const int lb = d.tlg < lc ? d.tlg : lc;
        const int run = L << lb;
        for (int cb = 0; cb < (C >> lb); ++cb) {
            for (int e = threadIdx.x; e < run; e += NT) {
                const int y = e >> lb;
                const int c = (cb << lb) + (e & ((1 << lb) - 1));
                const int gx = gx0 + c;
                if (gx < d.hxa) gstore(dst, t_at(d, gx, y), res[lp(c * L + y)]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// K2
// ---------------------------------------------------------------------------
// Physical fields of a lane's points.  `rho` holds the density after the
// inverse transforms and 1/rho (1/rho0 when incompressible) once the velocity
// is made; v = j * (1/rho) is formed where a product needs it, which keeps two
// arrays fewer in registers.
template <typename T, int PPT, int MAXN>
struct RowFields {
    T rho[PPT], jx[PPT], jy[PPT];
    T psi[MAXN > 0 ? MAXN : 1][PPT];

    __device__ __forceinline__ T vx(int i) const { return jx[i] * rho[i]; }
    __device__ __forceinline__ T vy(int i) const { return jy[i] * rho[i]; }
    // opaque to the optimizer: the velocity products of each output pass are
    // formed in that pass, not hoisted out of the pass loop and held live
    __device__ __forceinline__ void fence() {
#pragma unroll
        for (int i = 0; i < PPT; ++i) asm volatile("" : "+v"(rho[i]));
    }
    // rho -> 1/rho (compressible) or 1/rho0; true when the density is zero
    __device__ __forceinline__ bool make_velocity(int i, bool compressible, T r0) {
        const T r = compressible ? rho[i] : r0;
        if constexpr (sizeof(T) == 4) {
            // FP32 form: the hardware reciprocal (1 ulp) and one Newton step
            // instead of the correctly rounded division sequence (r = 0 gives
            // a non-finite value and the zero-density flag, as the division)
            const float e = __builtin_amdgcn_rcpf(r);
            rho[i] = __builtin_fmaf(e, __builtin_fmaf(-r, e, 1.0f), e);
        } else {
            rho[i] = T(1) / r;
        }
        return compressible && r == T(0);
    }

    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            rho[i] = jx[i] = jy[i] = T(0);
#pragma unroll
            for (int a = 0; a < (MAXN > 0 ? MAXN : 1); ++a) psi[a][i] = T(0);
        }
    }
    // i must be a compile-time constant after unrolling.  Every array is
    // rewritten through a select, never through a branch: branches on the
    // runtime role let the compiler merge the stores into one store through a
    // selected address, which puts the whole struct in scratch memory.
    __device__ __forceinline__ void store(int role, int i, T v) {
        rho[i] = role == ROLE_RHO ? v : rho[i];
        jx[i] = role == ROLE_JX ? v : jx[i];
        jy[i] = role == ROLE_JY ? v : jy[i];
#pragma unroll
        for (int a = 0; a < MAXN; ++a) psi[a][i] = role == ROLE_PSI0 + a ? v : psi[a][i];
    }
    // the same store for a role known at compile time
    template <int ROLE>
    __device__ __forceinline__ void store_ct(int i, T v) {
        if constexpr (ROLE == ROLE_RHO) rho[i] = v;
        else if constexpr (ROLE == ROLE_JX) jx[i] = v;
        else if constexpr (ROLE == ROLE_JY) jy[i] = v;
        else psi[ROLE - ROLE_PSI0][i] = v;
    }
    __device__ __forceinline__ T psi_of(int alpha, int i) const {
        T v = T(0);
#pragma unroll
        for (int a = 0; a < MAXN; ++a)
            if (alpha == a) v = psi[a][i];
        return v;
    }
    // f(psi[alpha]) for an order parameter index uniform across the wave:
    // with up to two fields one scalar branch picks the array (no select per
    // value); more fields keep the selects (one copy of the code).
    template <typename F>
    __device__ __forceinline__ void with_psi(int alpha, F&& f) const {
        if constexpr (MAXN == 2) {
            if (alpha == 0) f(psi[0]);
            else f(psi[1]);
        } else if constexpr (MAXN <= 1) {
            f(psi[0]);
        } else {
            T p[PPT];
#pragma unroll
            for (int i = 0; i < PPT; ++i) p[i] = psi_of(alpha, i);
            f(p);
        }
    }
    // Output `it` at every point of the lane: one branch on the output kind
    // per output, with the points in straight-line code inside each case
    // (a branch per point splits the arithmetic into blocks of a few
    // instructions each, which the scheduler cannot interleave).
    __device__ __forceinline__ void eval_all(const Desc& d, OutItem it, T (&out)[PPT]) const {
        switch (it.kind) {
            case OUT_MU: {
                const T u = T(d.phi4_u[it.alpha]);   // src/model_free_energy_phi4.cc:54
                with_psi(it.alpha, [&](const T (&p)[PPT]) {
#pragma unroll
                    for (int i = 0; i < PPT; ++i) out[i] = u * p[i] * p[i] * p[i];
                });
                break;
            }
            case OUT_PSIVX:
                with_psi(it.alpha, [&](const T (&p)[PPT]) {
#pragma unroll
                    for (int i = 0; i < PPT; ++i) out[i] = p[i] * vx(i);
                });
                break;
            case OUT_PSIVY:
                with_psi(it.alpha, [&](const T (&p)[PPT]) {
#pragma unroll
                    for (int i = 0; i < PPT; ++i) out[i] = p[i] * vy(i);
                });
                break;
            case OUT_GRADF: {
#pragma unroll
                for (int i = 0; i < PPT; ++i) out[i] = T(0);
                for (int g = 0; g < d.n_grad; ++g) {
                    if (d.grad_component[g] != it.alpha) continue;
                    const T a = T(-d.grad_amplitude[g]);
                    if (d.grad_direction[g] == 0) {
#pragma unroll
                        for (int i = 0; i < PPT; ++i) out[i] += a * vx(i);
                    } else {
#pragma unroll
                        for (int i = 0; i < PPT; ++i) out[i] += a * vy(i);
                    }
                }
                break;
            }
            case OUT_VX:
#pragma unroll
                for (int i = 0; i < PPT; ++i) out[i] = vx(i);
                break;
            case OUT_VY:
#pragma unroll
                for (int i = 0; i < PPT; ++i) out[i] = vy(i);
                break;
            case OUT_PXX:
#pragma unroll
                for (int i = 0; i < PPT; ++i) out[i] = jx[i] * vx(i);
                break;
            case OUT_PXY:
#pragma unroll
                for (int i = 0; i < PPT; ++i) out[i] = jx[i] * vy(i);
                break;
            default:
#pragma unroll
                for (int i = 0; i < PPT; ++i) out[i] = jy[i] * vy(i);
                break;
        }
    }
};

// Row transforms of K2.  FL > 0: a row kernel built for rows of exactly FL
// points, with only the fixed-length transform (no runtime-plan fallback
// competing for registers).  FL = 0: any length; the fixed-length dispatch in
// FP32, the runtime plan in FP64, where the dispatch costs the few registers
// that separate 3 from 2 resident waves per SIMD.
template <typename T, int NT, int PPT, int FL>
__device__ __forceinline__ void row_fft(cx<T>* buf, int L, int nseq, const FftPlanDev& plan,
                                        const cx<T>* __restrict__ tw, int sign,
                                        int tid = static_cast<int>(threadIdx.x)) {
    if constexpr (FL > 0) lds_fft_ct<T, NT, PPT, FL>(buf, nseq, tw, sign, tid);
    else if constexpr (sizeof(T) == 4) lds_fft_any<T, NT, PPT>(buf, L, nseq, plan, tw, sign, tid);
    else lds_fft<T, NT, PPT>(buf, L, nseq, plan, tw, sign, tid);
}

// Inverse-transformed fields in a fixed order, known at compile time:
// LAYOUT 1: psi_0 .. psi_{MAXN-1}, rho, j_x, j_y; 2: psi_0 .. psi_{MAXN-1},
// j_x, j_y; 3: psi_0 .. psi_{MAXN-1}; 0: any order (roles read at run time).
constexpr int layout_nin(int layout, int maxn) {
    return layout == 1 ? maxn + 3 : (layout == 2 ? maxn + 2 : maxn);
}
constexpr int layout_role(int layout, int maxn, int p) {
    if (p < maxn) return ROLE_PSI0 + p;
    if (layout == 1) return p == maxn ? ROLE_RHO : (p == maxn + 1 ? ROLE_JX : ROLE_JY);
    return p == maxn ? ROLE_JX : ROLE_JY;
}

// WS > 1 (the staged any-length build of short rows, where a launch has few
// workgroups): the workgroup's WS wave groups of NT / WS lanes hold the same
// R rows; group g inverse-transforms the field pairs g, g + WS, ... into its
// own LDS buffer and parks the physical fields in a shared LDS plane set,
// then every group reads all fields at its points and forward-transforms the
// product pairs g, g + WS, ...  Every group runs the same number of
// transforms (zeros where it has no pair left), so the barriers line up.
//
// NBUF > 0 (FL > 0 and PPT = 6 only): the inverse transform runs with the
// radix-3 pass last and leaves each lane the six points j, j + FL/3, j + 2FL/3
// of its two radix-3 butterflies; the products at those points feed the
// forward transform's radix-3 first pass straight from registers, the
// inverse's first pass loads T straight into its butterflies, and the passes
// alternate between NBUF LDS buffers.  NBUF = 0: every transform staged
// through one LDS buffer.
template <typename T, int NT, int PPT, int MAXN, int FL, int NBUF, int LAYOUT, int WS>
__global__ __launch_bounds__(NT) void k_rows(const InstDev<T>* __restrict__ insts, BlockMap map) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    // contiguous row ranges per XCD: the column-major U lines that
    // neighbouring rows fill meet in one L2
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int inst = map.inst[blk];
    const int local = blk - map.first[inst];
    const InstDev<T>& I = insts[inst];
    if (instance_failed(I.flag)) return;
    const Desc& d = I.d;
    const cx<T>* __restrict__ tbuf = I.tbuf;
    cx<T>* __restrict__ ubuf = I.ubuf;
    const cx<T>* __restrict__ twx = I.twx;
    int* __restrict__ flag = I.flag;
    // a fixed-length build knows its row length and rows per workgroup
    const int R = FL > 0 ? (NT * PPT) / FL : I.R;
    const int L = FL > 0 ? FL : d.mx;
    const int my = d.my;
    const int hxa = d.hxa;
    const int y0 = local * R;
    const int npts = R * L;
    const size_t plane = static_cast<size_t>(my) * hxa;
    const int tid = threadIdx.x;

    RowFields<T, PPT, MAXN> F;
    F.clear();

    if constexpr (WS > 1) {
        static_assert(FL == 0 && NBUF == 0 && LAYOUT == 0 && NT % WS == 0, "wave groups split the staged row pass");
        constexpr int NTF = NT / WS;
        const int grp = tid / NTF;
        const int tidf = tid - grp * NTF;
        const int lsz = lp_size(npts);   // one group's transform buffer (R L <= NTF PPT, host-checked)
        cx<T>* const gbuf = lds + static_cast<size_t>(grp) * lsz;
        T* const phys = reinterpret_cast<T*>(lds + static_cast<size_t>(WS) * lsz);   // n_in planes of R L
        // ---- inverse x transforms: field pair pr by group pr mod WS --------
        const int npairs = (d.n_in + 1) / 2;
        const int rounds_in = (npairs + WS - 1) / WS;
        for (int it = 0; it < rounds_in; ++it) {
            const int pr = it * WS + grp;
            const bool mine = pr < npairs;
            const int s = mine ? 2 * pr : 0;
            const bool has_b = mine && s + 1 < d.n_in;
            const cx<T>* __restrict__ ta = tbuf + static_cast<size_t>(s) * plane;
            const cx<T>* __restrict__ tb = tbuf + static_cast<size_t>(has_b ? s + 1 : s) * plane;
            cx<T> va[PPT], vb[PPT];
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int idx = tidf + NTF * i;
                const int r = idx / L;
                const int k = idx - r * L;
                const int y = y0 + r;
                const bool mirror = k > L - hxa;
                const int kk = mirror ? L - k : (k < hxa ? k : 0);
                const int yy = y < my ? y : 0;
                const uint32_t at = d.tlg < 0 ? static_cast<uint32_t>(yy * hxa + kk) : t_at(d, kk, yy);
                va[i] = mine ? gload(ta, at) : czero<T>();
                vb[i] = mine ? gload(tb, at) : czero<T>();
            }
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int idx = tidf + NTF * i;
                if (idx >= npts) continue;
                const int r = idx / L;
                const int k = idx - r * L;
                const bool mirror = k > L - hxa;
                const bool live = mine && y0 + r < my && (k < hxa || mirror);
                cx<T> a = va[i];
                cx<T> b = has_b ? vb[i] : czero<T>();
                a.im = k == 0 ? T(0) : (mirror ? -a.im : a.im);
                b.im = k == 0 ? T(0) : (mirror ? -b.im : b.im);
                gbuf[lp(idx)] = live ? mk(a.re - b.im, a.im + b.re) : czero<T>();   // A + iB
            }
            __syncthreads();
            row_fft<T, NTF, PPT, FL>(gbuf, L, R, d.plan_x, twx, +1, tidf);
            if (mine) {
#pragma unroll
                for (int i = 0; i < PPT; ++i) {
                    const int idx = tidf + NTF * i;
                    if (idx >= npts) continue;
                    const cx<T> z = gbuf[lp(idx)];
                    phys[s * npts + idx] = z.re;
                    if (has_b) phys[(s + 1) * npts + idx] = z.im;
                }
            }
            __syncthreads();
        }
        // every group: all fields at its points
        for (int f = 0; f < d.n_in; ++f) {
            const int role = d.in_role[f];
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int idx = tidf + NTF * i;
                F.store(role, i, idx < npts ? phys[f * npts + idx] : T(0));
            }
        }
        // ---- velocity (src/fcalculator_workspace.cc:122-133) ---------------
        if (d.make_velocity) {
            const T r0 = T(d.rho0);
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int idx = tidf + NTF * i;
                if (idx < npts && F.make_velocity(i, d.velocity_compressible != 0, r0) && y0 + idx / L < my)
                    atomicOr(flag, 1);
            }
        }
        // ---- products and forward x transforms: product pair q by group q mod WS
        const int nopairs = (d.n_out + 1) / 2;
        const int rounds_out = (nopairs + WS - 1) / WS;
        for (int it = 0; it < rounds_out; ++it) {
            const int q = it * WS + grp;
            const bool mine = q < nopairs;
            const int o = mine ? 2 * q : 0;
            const bool has_b = mine && o + 1 < d.n_out;
            const OutItem ia = d.out[o];
            const OutItem ib = has_b ? d.out[o + 1] : d.out[o];
            F.fence();
            T va[PPT], vb[PPT];
            F.eval_all(d, ia, va);
            if (has_b) F.eval_all(d, ib, vb);
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int idx = tidf + NTF * i;
                if (idx < npts) gbuf[lp(idx)] = mine ? mk(va[i], has_b ? vb[i] : T(0)) : czero<T>();
            }
            __syncthreads();
            row_fft<T, NTF, PPT, FL>(gbuf, L, R, d.plan_x, twx, -1, tidf);
            if (mine) {
                cx<T>* __restrict__ ua = ubuf + static_cast<size_t>(o) * plane;
                cx<T>* __restrict__ ub = ubuf + static_cast<size_t>(has_b ? o + 1 : o) * plane;
                const int nk = R * hxa;
                const int umask = (1 << d.ulg) - 1;
                for (int idx = tidf; idx < nk; idx += NTF) {
                    const int kb = idx / (R << d.ulg);
                    const int rem = idx - kb * (R << d.ulg);
                    const int r = rem >> d.ulg;
                    const int k = (kb << d.ulg) + (rem & umask);
                    const int y = y0 + r;
                    if (y < my) {
                        const cx<T> zk = gbuf[lp(r * L + k)];
                        const cx<T> zm = gbuf[lp(r * L + (k == 0 ? 0 : L - k))];
                        const uint32_t at = ((static_cast<uint32_t>(kb) * my + y) << d.ulg) + (k & umask);   // kx-blocked U
                        gstore(ua, at, mk(T(0.5) * (zk.re + zm.re), T(0.5) * (zk.im - zm.im)));
                        if (has_b) gstore(ub, at, mk(T(0.5) * (zk.im + zm.im), T(0.5) * (zm.re - zk.re)));
                    }
                }
            }
            __syncthreads();
        }
        return;
    }

    if constexpr (NBUF > 0) {
        static_assert(FL > 0 && PPT % 3 == 0, "fused row transforms need a fixed length and radix-3 butterflies per lane");
        constexpr int RF = NT * PPT / FL;              // rows per workgroup
        constexpr int R0 = fz_radix(FL, false, 0);
        using MI = FzMap<NT, FL, RF, R0>;
        constexpr int B3 = PPT / 3;                 // radix-3 butterflies per lane
        static_assert(FzMap<NT, FL, RF, 3>::NB == B3 * NT, "whole radix-3 butterflies per lane");
        // RF == 1: a launch has one workgroup per row of each instance, so
        // every workgroup's row is inside the grid (no row masks)
        // The inverse transforms' last radix-8 / radix-3 passes and the forward
        // transforms' radix-3 / first radix-8 passes exchange across lanes
        // (FzGrp map of the lane's six points) in FP64 below 1024 lanes.  The
        // 1024-lane FP64 builds are short of registers (they would spill
        // more), and FP32 keeps the LDS exchange: its 1536-point row pass is
        // bound by VALU issue, which the pair's fourth wave and transposes add
        // to (measured 22 % slower), and at 6144 points it measured 2 % slower.
        constexpr bool PAIR = sizeof(T) == 8 && NT < 1024 && PPT == 6 && fz_pair_ok<FL, false>() && fz_pair_ok<FL, true>();
        FzBufs<T, NBUF> bf{lds, lds + lp_size(RF * FL), 0};

        // ---- inverse x transforms, two fields per complex FFT -------------
        // fields s and s + 1 (if has_b) of T into w: lane b's butterfly holds
        // the points j, j + FL/3, j + 2FL/3 of its row
        auto transform_pair = [&](int s, bool has_b, cx<T> (&w)[B3][3]) {
            const cx<T>* __restrict__ ta = tbuf + static_cast<size_t>(s) * plane;
            const cx<T>* __restrict__ tb = tbuf + static_cast<size_t>(has_b ? s + 1 : s) * plane;
            cx<T> v[MI::B][R0];
            cx<T> vb[MI::B][R0];
            // an opaque copy of the lane index: the load offsets are formed
            // per transform, not held live across the loop
            int lane = tid;
            asm volatile("" : "+v"(lane));
            __builtin_assume(lane >= 0 && lane < NT);   // the range the copy hides
            // every load is issued before any is used, at an index that is
            // always inside the row (kx = 0 for the zero-padded points and
            // row 0 past the last row), and the padding is zeroed afterwards
            // by selects: a load under a per-point branch is waited for
            // inside its branch, which serializes the row's loads.  TH: the
            // 3/2-padded rows (Nax/2 = FL/3), where each of the lane's R0
            // points is known at compile time to be padding, kept or
            // mirrored, except the two runs that straddle FL/3 and 2FL/3.
            // BLK: kx-blocked T (Desc::tlg), else row-major.  LIN (TH and
            // BLK, 2^tlg dividing M): the kept point j + qM of a lane sits at
            // t_at(j) + qM My and the mirrored one FL - qM - j at
            // (FL - qM) My + t_at(-j), a lane base plus a uniform offset.
            auto load_points = [&]<bool TH, bool BLK, bool LIN>() {
                constexpr int M = MI::M;
                constexpr int H = FL / 3;
                const int hx = TH ? H : hxa;
#pragma unroll
                for (int b = 0; b < MI::B; ++b) {
                    const int id = lane + NT * b;
                    if (!MI::valid(id)) continue;
                    const int y = y0 + MI::seq(id);
                    const int j = MI::bfly(id);
                    const int yy = (RF == 1 || y < my) ? y : 0;
                    const uint32_t row = static_cast<uint32_t>(yy * hxa);
                    int kbase = 0, mbase = 0;
                    if constexpr (LIN) {
                        kbase = static_cast<int>(t_at(d, j, yy));
                        mbase = (((-j) >> d.tlg) * my << d.tlg) + (yy << d.tlg) + ((-j) & ((1 << d.tlg) - 1));
                    }
                    [&]<int... Q>(std::integer_sequence<int, Q...>) {
                        ([&] {
                            constexpr int q = Q;
                            constexpr bool dead = TH && q * M >= H && q * M + M - 1 <= FL - H;
                            constexpr bool kept = TH && q * M + M - 1 < H;           // all k < Nax/2
                            constexpr bool mirrored = TH && q * M > FL - H;          // all k > FL - Nax/2
                            if constexpr (!dead) {
                                // one load per field: kx = k (k < Nax/2) or the
                                // mirror kx = FL - k (k > FL - Nax/2) of the Hermitian row
                                const int k = j + q * M;
                                uint32_t at;
                                if constexpr (LIN && kept) {
                                    at = static_cast<uint32_t>(kbase + q * M * my);
                                } else if constexpr (LIN && mirrored) {
                                    at = static_cast<uint32_t>(mbase + (FL - q * M) * my);
                                } else {
                                    const bool mirror = mirrored || (!kept && k > FL - hx);
                                    const int kk = kept ? k : (mirror ? FL - k : (k < hx ? k : 0));
                                    at = BLK ? t_at(d, kk, yy) : row + kk;
                                }
                                v[b][q] = gload(ta, at);
                                vb[b][q] = gload(tb, at);   // tb = ta when there is no second field
                            }
                        }(), ...);
                    }(std::make_integer_sequence<int, R0>{});
                }
#pragma unroll
                for (int b = 0; b < MI::B; ++b) {
                    const int id = lane + NT * b;
                    if (!MI::valid(id)) continue;
                    const bool row_ok = RF == 1 || y0 + MI::seq(id) < my;
                    const int j = MI::bfly(id);
                    [&]<int... Q>(std::integer_sequence<int, Q...>) {
                        ([&] {
                            constexpr int q = Q;
                            constexpr bool dead = TH && q * M >= H && q * M + M - 1 <= FL - H;
                            constexpr bool kept = TH && q * M + M - 1 < H;           // all k < Nax/2
                            constexpr bool mirrored = TH && q * M > FL - H;          // all k > FL - Nax/2
                            if constexpr (dead) {
                                v[b][q] = czero<T>();
                            } else {
                                const int k = j + q * M;
                                const bool mirror = mirrored || (!kept && k > FL - hx);
                                const bool live = row_ok && (kept || mirrored || k < hx || mirror);
                                cx<T> a = v[b][q];
                                cx<T> bb = has_b ? vb[b][q] : czero<T>();
                                // the kx = 0 mode is real; the mirror half is conj(A) + i conj(B)
                                if constexpr (q == 0) {
                                    a.im = k == 0 ? T(0) : (mirror ? -a.im : a.im);
                                    bb.im = k == 0 ? T(0) : (mirror ? -bb.im : bb.im);
                                } else {
                                    a.im = mirror ? -a.im : a.im;
                                    bb.im = mirror ? -bb.im : bb.im;
                                }
                                const cx<T> z = mk(a.re - bb.im, a.im + bb.re);   // A + iB
                                v[b][q] = live ? z : czero<T>();
                            }
                        }(), ...);
                    }(std::make_integer_sequence<int, R0>{});
                }
            };
            // the pair's loads issue at raised wave priority, so they start
            // ahead of the other waves' transform arithmetic (measured: row
            // pass -0.6 % FP64, -0.8 % FP32 at 1024^2; the same around the
            // inverse- and forward-column loads measured slower)
            __builtin_amdgcn_s_setprio(2);
            if (d.tlg < 0) {
                if (3 * hxa == FL) load_points.template operator()<true, false, false>();
                else load_points.template operator()<false, false, false>();
            } else {
                if (3 * hxa == FL && (MI::M & ((1 << d.tlg) - 1)) == 0) load_points.template operator()<true, true, true>();
                else load_points.template operator()<false, true, false>();
            }
            __builtin_amdgcn_s_setprio(0);
            fz_first<T, NT, FL, RF, false, NBUF>(bf, v, +1, tid);
            if constexpr (PAIR) fz_rest_to_regs_pair<T, NT, FL, RF, NBUF>(bf, w, twx, +1, tid);
            else fz_rest_to_regs<T, NT, FL, RF, false, NBUF>(bf, w, twx, +1, tid);
        };
        // The last pass of each inverse transform is computed where it stands:
        // left to the optimizer, the butterflies of a field pair not needed
        // until the products sink past the later transforms, and their LDS
        // inputs and twiddles wait in scratch meanwhile.
        // (measured: 4096^2 FP64 row pass 2036 -> 1885 us; at 256 lanes, where
        // nothing spills, the fence costs 2 %, so it is kept to 1024 lanes)
        auto fence_outputs = [](cx<T> (&w)[B3][3]) {
            if constexpr (NT >= 1024) {
#pragma unroll
                for (int b = 0; b < B3; ++b)
#pragma unroll
                    for (int p = 0; p < 3; ++p) asm volatile("" : "+v"(w[b][p].re), "+v"(w[b][p].im));
            }
        };
        if constexpr (LAYOUT > 0) {
            // roles known at compile time: every value goes straight to its array
            constexpr int NIN = layout_nin(LAYOUT, MAXN);
            auto pair_ct = [&]<int S>() {
                constexpr bool HB = S + 1 < NIN;
                cx<T> w[B3][3];
                transform_pair(S, HB, w);
                fence_outputs(w);
#pragma unroll
                for (int b = 0; b < B3; ++b)
#pragma unroll
                    for (int p = 0; p < 3; ++p) {
                        F.template store_ct<layout_role(LAYOUT, MAXN, S)>(3 * b + p, w[b][p].re);
                        if constexpr (HB) F.template store_ct<layout_role(LAYOUT, MAXN, S + 1)>(3 * b + p, w[b][p].im);
                    }
            };
            [&]<int... P>(std::integer_sequence<int, P...>) {
                (pair_ct.template operator()<2 * P>(), ...);
            }(std::make_integer_sequence<int, (NIN + 1) / 2>{});
        } else {
            for (int s = 0; s < d.n_in; s += 2) {
                const bool has_b = s + 1 < d.n_in;
                cx<T> w[B3][3];
                transform_pair(s, has_b, w);
                fence_outputs(w);
                const int ra = d.in_role[s];
                const int rb = has_b ? d.in_role[s + 1] : -1;
#pragma unroll
                for (int b = 0; b < B3; ++b)
#pragma unroll
                    for (int p = 0; p < 3; ++p) {
                        F.store(ra, 3 * b + p, w[b][p].re);
                        if (rb >= 0) F.store(rb, 3 * b + p, w[b][p].im);
                    }
            }
        }

        // ---- velocity (src/fcalculator_workspace.cc:122-133) -----------------
        if (d.make_velocity) {
            const T r0 = T(d.rho0);
#pragma unroll
            for (int i = 0; i < PPT; ++i) {
                const int id = tid + NT * (i / 3);
                const int row = PAIR ? FzGrp<FL>(tid).s : FzMap<NT, FL, RF, 3>::seq(id);
                if (F.make_velocity(i, d.velocity_compressible != 0, r0) && (RF == 1 || y0 + row < my)) atomicOr(flag, 1);
            }
        }

        // ---- products and forward x transforms, two per complex FFT ---------
        for (int o = 0; o < d.n_out; o += 2) {
            const bool has_b = o + 1 < d.n_out;
            const OutItem ia = d.out[o];
            const OutItem ib = has_b ? d.out[o + 1] : d.out[o];
            F.fence();
            T va[PPT], vb[PPT];
            F.eval_all(d, ia, va);
            if (has_b) F.eval_all(d, ib, vb);
            else {
#pragma unroll
                for (int i = 0; i < PPT; ++i) vb[i] = T(0);
            }
            cx<T> v[B3][3];
#pragma unroll
            for (int b = 0; b < B3; ++b)
#pragma unroll
                for (int p = 0; p < 3; ++p) v[b][p] = mk(va[3 * b + p], vb[3 * b + p]);
            if constexpr (PAIR) {
                fz_first_pair<T, NT, FL, RF, NBUF>(bf, v, twx, -1, tid);
                fz_rest_to_lds_after_pair<T, NT, FL, RF, NBUF>(bf, twx, -1, tid);
            } else {
                fz_first<T, NT, FL, RF, true, NBUF>(bf, v, -1, tid);
                fz_rest_to_lds<T, NT, FL, RF, true, NBUF>(bf, twx, -1, tid);
            }
            __syncthreads();
            const cx<T>* res = bf.rd();
            cx<T>* __restrict__ ua = ubuf + static_cast<size_t>(o) * plane;
            cx<T>* __restrict__ ub = ubuf + static_cast<size_t>(has_b ? o + 1 : o) * plane;
            // U planes in kx blocks of 2^ulg columns over all rows: the lanes
            // of one block write one segment per row
            const int nk = RF * hxa;
            const int umask = (1 << d.ulg) - 1;
            static_assert((RF & (RF - 1)) == 0, "rows per workgroup a power of two");
            const int lgb = log2p2(RF) + d.ulg;      // log2 of the rows x columns of one U block segment
            if (sizeof(T) == 4 && RF == 1 && d.ulg >= 1 && (hxa & 1) == 0) {
                // FP32, one row (y0 < my) in blocks of at least two columns: a
                // lane takes the two columns kx = k, k + 1 (k even) of one block
                // segment (adjacent in U, one 16-byte store each plane; in FP64
                // the 32-byte stores touch twice the segments per instruction
                // and measured 20 % slower)
                const uint32_t yrow = static_cast<uint32_t>(y0) << d.ulg;
                for (int k = 2 * tid; k < hxa; k += 2 * NT) {
                    const cx<T> z0 = res[lp(k)];
                    const cx<T> z1 = res[lp(k + 1)];
                    const cx<T> m0 = res[lp(k == 0 ? 0 : FL - k)];
                    const cx<T> m1 = res[lp(FL - 1 - k)];
                    const uint32_t at = ((static_cast<uint32_t>(k >> d.ulg) * my) << d.ulg) + yrow + (k & umask);
                    gstore2(ua, at, mk(T(0.5) * (z0.re + m0.re), T(0.5) * (z0.im - m0.im)),
                            mk(T(0.5) * (z1.re + m1.re), T(0.5) * (z1.im - m1.im)));
                    if (has_b)
                        gstore2(ub, at, mk(T(0.5) * (z0.im + m0.im), T(0.5) * (m0.re - z0.re)),
                                mk(T(0.5) * (z1.im + m1.im), T(0.5) * (m1.re - z1.re)));
                }
                continue;
            }
            for (int idx = tid; idx < nk; idx += NT) {
                const int kb = idx >> lgb;
                const int rem = idx - (kb << lgb);
                const int r = rem >> d.ulg;
                const int k = (kb << d.ulg) + (rem & umask);
                const int y = y0 + r;
                if (y < my) {
                    const cx<T> zk = res[lp(r * FL + k)];
                    const cx<T> zm = res[lp(r * FL + (k == 0 ? 0 : FL - k))];
                    const uint32_t at = ((static_cast<uint32_t>(kb) * my + y) << d.ulg) + (k & umask);
                    gstore(ua, at, mk(T(0.5) * (zk.re + zm.re), T(0.5) * (zk.im - zm.im)));
                    if (has_b) gstore(ub, at, mk(T(0.5) * (zk.im + zm.im), T(0.5) * (zm.re - zk.re)));
                }
            }
        }
        return;
    }

    // ---- inverse x transforms, two fields per complex FFT -----------------
    for (int s = 0; s < d.n_in; s += 2) {
        const bool has_b = s + 1 < d.n_in;
        const cx<T>* __restrict__ ta = tbuf + static_cast<size_t>(s) * plane;
        const cx<T>* __restrict__ tb = tbuf + static_cast<size_t>(has_b ? s + 1 : s) * plane;
        // point idx = r L + k of the rows (R L <= NT PPT, host-checked); all
        // loads issued first at indices inside the rows (kx = 0 for the
        // zero-padded points, row 0 past the last row), masks applied after
        cx<T> va[PPT], vb[PPT];
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const int idx = tid + NT * i;
            const int r = idx / L;
            const int k = idx - r * L;
            const int y = y0 + r;
            const bool mirror = k > L - hxa;
            const int kk = mirror ? L - k : (k < hxa ? k : 0);
            const int yy = y < my ? y : 0;
            const uint32_t at = d.tlg < 0 ? static_cast<uint32_t>(yy * hxa + kk) : t_at(d, kk, yy);
            va[i] = gload(ta, at);
            vb[i] = gload(tb, at);   // tb = ta when there is no second field
        }
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const int idx = tid + NT * i;
            if (idx >= npts) continue;
            const int r = idx / L;
            const int k = idx - r * L;
            const bool mirror = k > L - hxa;
            const bool live = y0 + r < my && (k < hxa || mirror);
            cx<T> a = va[i];
            cx<T> b = has_b ? vb[i] : czero<T>();
            // the kx = 0 mode is real; the mirror half is conj(A) + i conj(B)
            a.im = k == 0 ? T(0) : (mirror ? -a.im : a.im);
            b.im = k == 0 ? T(0) : (mirror ? -b.im : b.im);
            lds[lp(idx)] = live ? mk(a.re - b.im, a.im + b.re) : czero<T>();   // A + iB
        }
        __syncthreads();
        row_fft<T, NT, PPT, FL>(lds, L, R, d.plan_x, twx, +1);
        const int ra = d.in_role[s];
        const int rb = has_b ? d.in_role[s + 1] : -1;
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const int idx = tid + NT * i;
            if (idx < npts) {
                const cx<T> z = lds[lp(idx)];
                F.store(ra, i, z.re);
                if (rb >= 0) F.store(rb, i, z.im);
            }
        }
        __syncthreads();
    }

    // ---- velocity (src/fcalculator_workspace.cc:122-133) ---------------------
    if (d.make_velocity) {
        const T r0 = T(d.rho0);
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const int idx = tid + NT * i;
            if (idx < npts && F.make_velocity(i, d.velocity_compressible != 0, r0) && y0 + idx / L < my)
                atomicOr(flag, 1);
        }
    }

    // ---- products and forward x transforms, two per complex FFT -------------
    for (int o = 0; o < d.n_out; o += 2) {
        const bool has_b = o + 1 < d.n_out;
        const OutItem ia = d.out[o];
        const OutItem ib = has_b ? d.out[o + 1] : d.out[o];
        F.fence();
        T va[PPT], vb[PPT];
        F.eval_all(d, ia, va);
        if (has_b) F.eval_all(d, ib, vb);
        else {
#pragma unroll
            for (int i = 0; i < PPT; ++i) vb[i] = T(0);
        }
#pragma unroll
        for (int i = 0; i < PPT; ++i) {
            const int idx = tid + NT * i;
            if (idx < npts) lds[lp(idx)] = mk(va[i], vb[i]);
        }
        __syncthreads();
        row_fft<T, NT, PPT, FL>(lds, L, R, d.plan_x, twx, -1);
        cx<T>* __restrict__ ua = ubuf + static_cast<size_t>(o) * plane;
        cx<T>* __restrict__ ub = ubuf + static_cast<size_t>(has_b ? o + 1 : o) * plane;
        const int nk = R * hxa;
        const int umask = (1 << d.ulg) - 1;
        for (int idx = tid; idx < nk; idx += NT) {
            const int kb = idx / (R << d.ulg);
            const int rem = idx - kb * (R << d.ulg);
            const int r = rem >> d.ulg;
            const int k = (kb << d.ulg) + (rem & umask);
            const int y = y0 + r;
            if (y < my) {
                const cx<T> zk = lds[lp(r * L + k)];
                const cx<T> zm = lds[lp(r * L + (k == 0 ? 0 : L - k))];
                const uint32_t at = ((static_cast<uint32_t>(kb) * my + y) << d.ulg) + (k & umask);   // kx-blocked U
                gstore(ua, at, mk(T(0.5) * (zk.re + zm.re), T(0.5) * (zk.im - zm.im)));
                if (has_b) gstore(ub, at, mk(T(0.5) * (zk.im + zm.im), T(0.5) * (zm.re - zk.re)));
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Physical fields for a snapshot (PhysicalStateBuffer of the reference,
// src/buffer_physical_state.h, filled by FourierTransform2D::backward_many):
// the row half of the inverse transform of every field, after K1 has done the
// column half into T.  Same pairing of two real fields per complex FFT as
// K2; rows are written to out[f][y][0..Mx).
// ---------------------------------------------------------------------------
template <typename T, int NT, int PPT>
__global__ __launch_bounds__(NT) void k_rows_physical(const InstDev<T>* __restrict__ insts, T* __restrict__ out,
                                                      int nfields) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const InstDev<T>& I = insts[0];
    const Desc& d = I.d;
    const cx<T>* __restrict__ tbuf = I.tbuf;
    const int R = I.R;
    const int L = d.mx;
    const int my = d.my;
    const int hxa = d.hxa;
    const int y0 = blockIdx.x * R;
    const int npts = R * L;
    const size_t plane = static_cast<size_t>(my) * hxa;
    for (int s = 0; s < nfields; s += 2) {
        const bool has_b = s + 1 < nfields;
        const cx<T>* __restrict__ ta = tbuf + static_cast<size_t>(s) * plane;
        const cx<T>* __restrict__ tb = tbuf + static_cast<size_t>(has_b ? s + 1 : s) * plane;
        for (int idx = threadIdx.x; idx < npts; idx += NT) {
            const int r = idx / L;
            const int k = idx - r * L;
            const int y = y0 + r;
            cx<T> z = czero<T>();
            if (y < my) {
                auto at = [&](int kx) {
                    return d.tlg < 0 ? static_cast<uint32_t>(y * hxa + kx) : t_at(d, kx, y);
                };
                if (k < hxa) {
                    cx<T> a = gload(ta, at(k));
                    cx<T> b = has_b ? gload(tb, at(k)) : czero<T>();
                    if (k == 0) { a.im = T(0); b.im = T(0); }
                    z = mk(a.re - b.im, a.im + b.re);
                } else if (k > L - hxa) {
                    const int kk = L - k;
                    const cx<T> a = gload(ta, at(kk));
                    const cx<T> b = has_b ? gload(tb, at(kk)) : czero<T>();
                    z = mk(a.re + b.im, b.re - a.im);
                }
            }
            lds[lp(idx)] = z;
        }
        __syncthreads();
        lds_fft_any<T, NT, PPT>(lds, L, R, d.plan_x, I.twx, +1);
        for (int idx = threadIdx.x; idx < npts; idx += NT) {
            const int r = idx / L;
            const int x = idx - r * L;
            const int y = y0 + r;
            if (y < my) {
                const cx<T> z = lds[lp(idx)];
                out[(static_cast<size_t>(s) * my + y) * L + x] = z.re;
                if (has_b) out[(static_cast<size_t>(s + 1) * my + y) * L + x] = z.im;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// noise: two complex N(0,1) deviates per Philox call.  The four 32-bit words
// give two Box-Muller pairs, each from a radius uniform u1 = (r + 1/2) 2^-32
// (so the tails reach sqrt(-2 ln 2^-33) = 6.8 sigma) and an angle uniform
// u2 = r' 2^-32.  The transcendentals are the hardware single-precision
// instructions in both arithmetic types: base-2 logarithm (u1 lies in
// [2^-33, 1], never denormal), square root, and sine and cosine, which take
// their argument in revolutions, so sin(2 pi u2) is one instruction on u2.
// The deviates agree with a double-precision evaluation of the same formula
// to ~1e-6, far below any sampling error a run can resolve.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void gauss_quad(const Desc& d, uint32_t mode_id, uint32_t stream, const StageParams& sp,
                                           cx<T>& z0, cx<T>& z1) {
    u32x4 ctr;
    ctr.v[0] = mode_id;
    ctr.v[1] = stream << 4;
    ctr.v[2] = sp.step_lo;
    ctr.v[3] = sp.step_hi;
    const u32x4 r = philox4x32_10(ctr, d.seed_lo, d.seed_hi);
    auto box_muller = [](uint32_t a, uint32_t b) {
        constexpr float kMinus2Ln2 = -1.38629436111989061883f;   // -2 ln 2: -2 ln u = -2 ln2 log2 u
        const float u1 = (static_cast<float>(a) + 0.5f) * 0x1.0p-32f;
        const float u2 = static_cast<float>(b) * 0x1.0p-32f;
        const float rad = __builtin_amdgcn_sqrtf(kMinus2Ln2 * __builtin_amdgcn_logf(u1));
        const float sn = __builtin_amdgcn_sinf(u2);
        const float cs = __builtin_amdgcn_cosf(u2);
        return mk(static_cast<T>(rad * cs), static_cast<T>(rad * sn));
    };
    z0 = box_muller(r.v[0], r.v[1]);
    z1 = box_muller(r.v[2], r.v[3]);
}

// FP32 form: the k = 0 mode is advanced in FP64 from its FP64 copies; its
// increment dt*det (noise vanishes at k = 0) comes from the FP32 pipeline and
// the self-conjugate mode stays real (src/time_integrator.h:85-88).
template <typename T>
__device__ __forceinline__ void k0_write(const cx<double>* base_k0, const cx<double>* cur_k0, cx<double>* next_k0,
                                         cx<T>* dst, int f, cx<T> inc, const StageParams& sp) {
    cx<double> nk;
    nk.re = sp.wb * base_k0[f].re + sp.wc * (cur_k0[f].re + static_cast<double>(inc.re));
    nk.im = 0.0;
    next_k0[f] = nk;
    *dst = mk(static_cast<T>(nk.re), T(0));
}

// ---------------------------------------------------------------------------
// K3
// ---------------------------------------------------------------------------
// FZL > 0: a build for columns of exactly FZL points, C = NT * VPT / FZL per
// workgroup, with Nay <= 2 FZL / 3.  The transform loads U straight into its
// first butterflies and ends with the radix-3 pass in registers: lane b's
// butterfly j (of a column s) then holds ky = j (output j) and ky = j - FZL/3
// (output j + 2FZL/3), the only kept outputs, so every lane accumulates two
// modes per butterfly and no LDS read-back is needed.  The passes alternate
// between NBUF LDS buffers.
//
// WS > 1 (small columns, few workgroups): the workgroup's WS wave groups of
// NT / WS lanes each transform their own products of the same column tile
// (product i by group i mod WS) with their own LDS buffers, park their
// accumulated terms side by side, and the stage update, spread over all NT
// lanes, adds the WS partial sums of each mode.  Every group runs the same
// number of transforms (a group out of products transforms zeros and keeps
// nothing), so the barriers inside the transforms line up.  While the step
// flux is recorded, group 0 takes every product (the flux slots of a mode
// are then added to in one order, by one lane).
// FP32 fused builds of 256 lanes are held to 80 VGPRs (six waves per SIMD
// instead of five at 83, no spills): a 1024^2 launch (1536 workgroups) then
// fits the GPU in one round instead of 1.2 (measured 44.2 -> 37.3 us).
template <typename T, int NT, int VPT, int ACC, int FZL, int NBUF, int WS>
constexpr int k3_min_waves() { return (sizeof(T) == 4 && NT == 256 && FZL > 0 && WS == 1) ? 6 : 1; }
template <typename T, int NT, int VPT, int ACC, int FZL, int NBUF, int WS>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(k3_min_waves<T, NT, VPT, ACC, FZL, NBUF, WS>(), 8)))
void k_forward_columns(const InstDev<T>* __restrict__ insts,
                                                        const StagePtrs<T>* __restrict__ ptrs, StageParams sp,
                                                        BlockMap map) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem_raw);
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int inst = map.inst[blk];
    int local = blk - map.first[inst];
    const InstDev<T>& I = insts[inst];
    if (instance_failed(I.flag)) return;
    const Desc& d = I.d;
    const StagePtrs<T>& P = ptrs[inst];
    const int C = I.C3;
    // Workgroup order (the momentum group has five transforms and three
    // fields to the psi groups' three and one; the dispatcher hands an XCD's
    // workgroups out in order).  A launch of a few rounds of resident
    // workgroups with the instance alone in it takes each XCD's range of
    // slots momentum groups first, so the long ones start first and the short
    // ones fill in behind them; a longer launch keeps the groups mixed (both
    // the memory and the arithmetic pipes stay busy) with the momentum group
    // first in each tile, so no range ends on a long workgroup.
    if (d.ngroups > 1 && d.groups[d.ngroups - 1].kind == GROUP_MOM) {
        if (I.k3_order == 1 && I.ntiles3 * d.ngroups == static_cast<int>(gridDim.x)) {
            constexpr int X = 8;
            const int b = blockIdx.x, nb = gridDim.x;
            const int x = b % X, q = nb / X, r = nb % X;
            local = heavy_first(x * q + (x < r ? x : r), q + (x < r ? 1 : 0), b / X, d.ngroups);
        } else {
            const int t = local / d.ngroups, g = local - t * d.ngroups;
            local = t * d.ngroups + (g == 0 ? d.ngroups - 1 : g - 1);
        }
    }
    const cx<T>* __restrict__ ubuf = I.ubuf;
    const cx<T>* __restrict__ twy = I.twy;
    {
        const uint64_t st = *I.step;
        sp.step_lo = static_cast<uint32_t>(st);
        sp.step_hi = static_cast<uint32_t>(st >> 32);
    }
    // tile-major order: every XCD's range of workgroups holds all groups of its
    // tiles, so the heavier momentum group does not pile up on a few XCDs
    const int tile = local / d.ngroups;
    const int group = local - tile * d.ngroups;
    const Group& g = d.groups[group];
    const int L = d.my;
    const int nay = d.nay;
    const int half = nay / 2;
    const int hxa = d.hxa;
    const int gx0 = tile * C;
    const int nmodes = C * nay;
    const size_t plane = static_cast<size_t>(L) * hxa;
    const int tid = threadIdx.x;

    // ---- flux slots this group records (FluxBuffer, src/buffer_flux.h) -------
    // Every mode is read and written by the one lane that owns it in all the
    // loops below (m = tid + NT * i), so the accumulation needs no atomics.
    const size_t fe = static_cast<size_t>(hxa) * nay;
    cx<T>* fx_slot = nullptr;   // psi group: psi flux x; momentum group: rho flux x
    cx<T>* fy_slot = nullptr;
    cx<T>* pi_slot = nullptr;   // Pi xx, then Pi xy and Pi yy at +fe, +2fe
    if (I.flux) {
        if (g.kind == GROUP_PSI) {
            if (I.flux_req & FLUX_ORDER_PARAMETER) {
                fx_slot = I.flux + static_cast<size_t>(2 + 2 * g.alpha) * fe;
                fy_slot = fx_slot + fe;
            }
        } else {
            if ((I.flux_req & FLUX_DENSITY) && d.mode == SHHD_MODE_COMPRESSIBLE) {
                fx_slot = I.flux;
                fy_slot = I.flux + fe;
            }
            if (I.flux_req & FLUX_MOMENTUM) pi_slot = I.flux + static_cast<size_t>(2 + 2 * d.N) * fe;
        }
        if (sp.flux_first) {   // FluxBuffer::begin_step
            for (int m = tid; m < nmodes; m += NT) {
                const int c = divn(m, nay, d.lnay);
                const int gx = gx0 + c;
                if (gx >= hxa) continue;
                const size_t at = static_cast<size_t>(gx) * nay + (m - c * nay);
                if (fx_slot) { fx_slot[at] = czero<T>(); fy_slot[at] = czero<T>(); }
                if (pi_slot) { pi_slot[at] = czero<T>(); pi_slot[at + fe] = czero<T>(); pi_slot[at + 2 * fe] = czero<T>(); }
            }
        }
    }
    const T fw = T(sp.flux_w);

    // ACC >= ceil(C * Nay / NT) accumulator slots per lane (host-checked)
    cx<T> acc0[ACC], acc1[ACC];
#pragma unroll
    for (int i = 0; i < ACC; ++i) { acc0[i] = czero<T>(); acc1[i] = czero<T>(); }

    // ---- forward y transforms of the products feeding this group ------------
    const bool psi_group = g.kind == GROUP_PSI;
    const bool recording = psi_group ? fx_slot != nullptr : pi_slot != nullptr;
    // one transformed mode v of column gx, signed ky_i, compact slot cidx
    // into accumulator slot i (and the step flux when recording)
    // Only the coefficients a product has (cf.nz, a uniform mask) are
    // evaluated: the branches are scalar and skip the vector work of the
    // zero terms.
    auto accumulate = [&](const ContribCoef& cf, int i, cx<T> v, int gx, int ky_i, int cidx) {
        const unsigned nz = cf.nz;
        const T kx = T(d.dkx) * T(gx);
        const T ky = T(d.dky) * T(ky_i);
        const cx<T> mv = mul_mi(v);   // -i v
        // a[r][0] + a[r][1] kx^2 + a[r][2] ky^2 + a[r][3] kx ky and b[r][0] kx + b[r][1] ky
        auto poly = [&](int r) {
            const unsigned m = nz >> (4 * r);
            T p = T(0);
            if (m & 1u) p = p + coef_a<T>(cf, r, 0);
            if (m & 2u) p = p + coef_a<T>(cf, r, 1) * (kx * kx);
            if (m & 4u) p = p + coef_a<T>(cf, r, 2) * (ky * ky);
            if (m & 8u) p = p + coef_a<T>(cf, r, 3) * (kx * ky);
            return p;
        };
        auto lin = [&](int r) {
            const unsigned m = nz >> (8 + 2 * r);
            T p = T(0);
            if (m & 1u) p = p + coef_b<T>(cf, r, 0) * kx;
            if (m & 2u) p = p + coef_b<T>(cf, r, 1) * ky;
            return p;
        };
        if (nz & 0x00Fu) acc0[i] = acc0[i] + poly(0) * v;
        if (nz & 0x300u) acc0[i] = acc0[i] + lin(0) * mv;
        if (!psi_group) {
            if (nz & 0x0F0u) acc1[i] = acc1[i] + poly(1) * v;
            if (nz & 0xC00u) acc1[i] = acc1[i] + lin(1) * mv;
        }
        if (recording) {
            const size_t at = static_cast<size_t>(gx) * nay + cidx;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
                cx<T>* dst = psi_group ? (t == 0 ? fx_slot : (t == 1 ? fy_slot : nullptr)) : pi_slot + t * fe;
                if (dst == nullptr || ((nz >> (12 + 3 * t)) & 7u) == 0) continue;
                const cx<T> fv = coef_f<T>(cf, t, 0) * v + (coef_f<T>(cf, t, 1) * kx + coef_f<T>(cf, t, 2) * ky) * mv;
                dst[at] = dst[at] + fw * fv;
            }
        }
    };

    // WS partial sums per mode, side by side (WS = 1: the sums themselves)
    cx<T>* acc_x = lds;
    cx<T>* acc_y = lds + WS * nmodes;
    static_assert(WS == 1 || FZL > 0, "wave groups split the fused column transforms only");
    if constexpr (FZL > 0) {
        constexpr int NTF = NT / WS;                // lanes of one transform group
        constexpr int CF = NTF * VPT / FZL;
        constexpr int M3 = FZL / 3;
        constexpr int R0 = fz_radix(FZL, false, 0);
        constexpr int NP = fz_npass(FZL, false);
        constexpr int B3 = VPT / 3;                 // radix-3 butterflies per lane
        static_assert(fz_radix(FZL, false, NP - 1) == 3 && ACC == 2 * B3, "radix-3 last pass, two modes per butterfly");
        static_assert(FzMap<NTF, FZL, CF, 3>::NB == B3 * NTF, "whole radix-3 butterflies per lane");
        static_assert(NT % WS == 0 && NTF % 64 == 0, "whole waves per group");
        using M0 = FzMap<NTF, FZL, CF, R0>;
        const int grp = WS > 1 ? tid / NTF : 0;     // transform group of this lane
        const int tidf = tid - grp * NTF;           // lane within the group
        // The last radix-8 and the radix-3 pass exchange across lanes (FP64
        // below 1024 lanes, as the row pass); the kept modes then follow the
        // FzGrp map
        constexpr bool PAIR = sizeof(T) == 8 && NTF < 1024 && VPT == 6 && fz_pair_ok<FZL, false>();
        auto seq_of = [&](int b, int lane_) { return PAIR ? FzGrp<FZL>(lane_).s : FzMap<NTF, FZL, CF, 3>::seq(lane_ + NTF * b); };
        auto bfly_of = [&](int b, int lane_) {
            return PAIR ? FzGrp<FZL>(lane_).j3(b) : FzMap<NTF, FZL, CF, 3>::bfly(lane_ + NTF * b);
        };
        cx<T>* const gbuf = lds + static_cast<size_t>(grp) * NBUF * lp_size(CF * FZL);
        FzBufs<T, NBUF> bf{gbuf, gbuf + lp_size(CF * FZL), 0};
        // the first butterflies of contribution ci, straight from U (zeros for
        // ci = none: a group with no product left in this iteration)
        const int ncon = g.ncontrib;
        auto load_contrib = [&](int ci, cx<T> (&v)[M0::B][R0]) {
            if (ci >= ncon) {
#pragma unroll
                for (int b = 0; b < M0::B; ++b)
#pragma unroll
                    for (int q = 0; q < R0; ++q) v[b][q] = czero<T>();
                return;
            }
            const cx<T>* __restrict__ src = ubuf + static_cast<size_t>(g.contrib[ci]) * plane;
            int lane0 = tidf;
            asm volatile("" : "+v"(lane0));
            __builtin_assume(lane0 >= 0 && lane0 < NTF);   // the range the copy hides
#pragma unroll
            for (int b = 0; b < M0::B; ++b) {
                const int id = lane0 + NTF * b;
                if (!M0::valid(id)) continue;
                const int s = M0::seq(id);
                const int j = M0::bfly(id);
                const int gx = gx0 + s;
                const uint32_t col = (static_cast<uint32_t>(gx >> d.ulg) * L << d.ulg) + (gx & ((1 << d.ulg) - 1));
#pragma unroll
                for (int q = 0; q < R0; ++q)   // kx-blocked U: consecutive lanes, consecutive rows
                    v[b][q] = gx < hxa ? gload(src, col + (static_cast<uint32_t>(j + q * M0::M) << d.ulg)) : czero<T>();
            }
        };
        // round it of group grp transforms product ci_of(it) (ncon: none)
        const bool split = WS > 1 && !recording;
        const int rounds = split ? (ncon + WS - 1) / WS : ncon;
        auto ci_of = [&](int it) { return split ? it * WS + grp : (grp == 0 ? it : ncon); };
        cx<T> v[M0::B][R0];
        if (rounds > 0) load_contrib(ci_of(0), v);
        for (int it = 0; it < rounds; ++it) {
            const int ci = ci_of(it);
            const ContribCoef& cf = g.coef[ci < ncon ? ci : 0];
            fz_first<T, NTF, FZL, CF, false, NBUF>(bf, v, -1, tidf);
            // the next contribution's loads fly while this one's passes run
            // (the barriers wait for LDS operations only)
            if (it + 1 < rounds) load_contrib(ci_of(it + 1), v);
            cx<T> w[B3][3];
            if constexpr (PAIR) fz_rest_to_regs_pair<T, NTF, FZL, CF, NBUF>(bf, w, twy, -1, tidf);
            else fz_rest_to_regs<T, NTF, FZL, CF, false, NBUF>(bf, w, twy, -1, tidf);
            if (ci >= ncon) continue;
            // an opaque copy of the lane index: the mode coordinates below are
            // recomputed per contribution instead of held live across the transform
            int lane = tidf;
            asm volatile("" : "+v"(lane));
            __builtin_assume(lane >= 0 && lane < NTF);   // the range the copy hides
#pragma unroll
            for (int b = 0; b < B3; ++b) {
                const int s = seq_of(b, lane);
                const int j = bfly_of(b, lane);
                const int gx = gx0 + s;
                if (gx >= hxa) continue;
                if (j < half) accumulate(cf, 2 * b, w[b][0], gx, j, j);
                if (M3 - j < half) accumulate(cf, 2 * b + 1, w[b][2], gx, j - M3, j - M3 + nay);
            }

        }
        // park the accumulated terms at their modes once every lane is done
        // with the transform buffers (group grp's sums at grp * nmodes)
        __syncthreads();
        cx<T>* const px = acc_x + grp * nmodes;
        cx<T>* const py = acc_y + grp * nmodes;
#pragma unroll
        for (int b = 0; b < B3; ++b) {
            const int s = seq_of(b, tidf);
            const int j = bfly_of(b, tidf);
            if (gx0 + s >= hxa) continue;
            if (j < half) {
                px[s * nay + j] = acc0[2 * b];
                py[s * nay + j] = acc1[2 * b];
            }
            if (M3 - j < half) {
                px[s * nay + j - M3 + nay] = acc0[2 * b + 1];
                py[s * nay + j - M3 + nay] = acc1[2 * b + 1];
            }
        }
        __syncthreads();
    } else {
        for (int ci = 0; ci < g.ncontrib; ++ci) {
            const int u = g.contrib[ci];
            const ContribCoef& cf = g.coef[ci];
            const cx<T>* __restrict__ src = ubuf + static_cast<size_t>(u) * plane;
            for (int c = 0; c < C; ++c) {   // kx-blocked U: a column's rows 2^ulg apart
                const int gx = gx0 + c;
                const uint32_t col = (static_cast<uint32_t>(gx >> d.ulg) * L << d.ulg) + (gx & ((1 << d.ulg) - 1));
                for (int y = tid; y < L; y += NT)
                    lds[lp(c * L + y)] = gx < hxa ? gload(src, col + (static_cast<uint32_t>(y) << d.ulg)) : czero<T>();
            }
            __syncthreads();
            col_fft<T, NT, VPT>(lds, L, C, d, twy, -1);
#pragma unroll
            for (int i = 0; i < ACC; ++i) {
                const int m = tid + NT * i;
                if (m >= nmodes) continue;
                const int c = divn(m, nay, d.lnay);
                const int cidx = m - c * nay;
                const int gx = gx0 + c;
                if (gx >= hxa || cidx == half) continue;
                const int ky_i = cidx < half ? cidx : cidx - nay;
                const int gy = ky_i >= 0 ? ky_i : ky_i + L;
                accumulate(cf, i, lds[lp(c * L + gy)], gx, ky_i, cidx);
            }
            __syncthreads();
        }

        // ---- park the accumulated product terms in LDS (free after the last
        // barrier) so the stage update below can walk the modes in a plain loop
#pragma unroll
        for (int i = 0; i < ACC; ++i) {
            const int m = tid + NT * i;
            if (m < nmodes) {
                acc_x[m] = acc0[i];
                acc_y[m] = acc1[i];
            }
        }
        __syncthreads();
    }

    // ---- spectral terms, noise, stage update --------------------------------
    // the product terms of mode m: the sum of the WS groups' parked sums
    auto acc_sum = [&](const cx<T>* acc, int m) {
        cx<T> r = acc[m];
#pragma unroll
        for (int w = 1; w < WS; ++w) r = r + acc[w * nmodes + m];
        return r;
    };
    const int a = g.alpha;
    const bool compressible = d.mode == SHHD_MODE_COMPRESSIBLE;
    const int fx_field = d.N + 1, fy_field = d.N + 2;
    T dt = T(d.dt), sdt = T(d.sqrt_dt);
    T wb = T(sp.wb), wc = T(sp.wc), beta = T(sp.beta);
    T dkx = T(d.dkx), dky = T(d.dky);
    // held in vector registers: the stage update's pointers, flags and
    // coefficients exceed the scalar registers, and these values are read by
    // vector instructions only
    asm volatile("" : "+v"(dt), "+v"(sdt), "+v"(wb), "+v"(wc), "+v"(beta), "+v"(dkx), "+v"(dky));
    constexpr bool kFp32 = sizeof(T) == 4;

    // one loop per group kind, chosen once per workgroup: each loop holds only
    // the pointers and coefficients of its own fields
    auto stage_modes = [&](auto psi_tag) {
        constexpr bool kPsi = decltype(psi_tag)::value;
        // The state loads of PF modes of the lane (FP64 below 1024 lanes:
        // two; the modes m = tid + NT i, i < ACC, host-checked) are issued
        // together before any of them is updated; a slot the lane does not
        // own loads mode 0 of the field and is not used.  The base state is
        // read only at the stages that weight it.  (FP32 and the 1024-lane
        // builds, capped at 128 registers, keep one mode at a time: the extra
        // registers measured slower there.)
        constexpr int NF = kPsi ? 1 : 3;   // psi_a | rho, jx, jy
        constexpr int PF = (sizeof(T) == 8 && NT < 1024 && ACC >= 2) ? 2 : 1;
        const bool use_base = sp.wb != 0.0;
        const bool need_rho = !kPsi && (compressible || (d.mom_det_on && d.pressure_coefficient != 0.0));
        auto field_of = [&](int q) { return kPsi ? 1 + a : (q == 0 ? 0 : (q == 1 ? fx_field : fy_field)); };
        for (int i0 = 0; i0 < ACC && tid + NT * i0 < nmodes; i0 += PF) {
            cx<T> pcur[PF][NF], pbas[PF][NF];
#pragma unroll
            for (int ii = 0; ii < PF; ++ii) {
                const int m = tid + NT * (i0 + ii);
                const int c = divn(m, nay, d.lnay);
                const bool own = i0 + ii < ACC && m < nmodes && gx0 + c < hxa;
                const uint32_t at = own ? static_cast<uint32_t>((gx0 + c) * nay + (m - c * nay)) : 0u;
#pragma unroll
                for (int q = 0; q < NF; ++q) {
                    pcur[ii][q] = (q > 0 || kPsi || need_rho) ? gload(P.cur[field_of(q)], at) : czero<T>();
                    pbas[ii][q] = czero<T>();
                }
            }
            if (use_base) {
#pragma unroll
                for (int ii = 0; ii < PF; ++ii) {
                    const int m = tid + NT * (i0 + ii);
                    const int c = divn(m, nay, d.lnay);
                    const bool own = i0 + ii < ACC && m < nmodes && gx0 + c < hxa;
                    const uint32_t at = own ? static_cast<uint32_t>((gx0 + c) * nay + (m - c * nay)) : 0u;
#pragma unroll
                    for (int q = 0; q < NF; ++q)
                        if (q > 0 || kPsi || compressible) pbas[ii][q] = gload(P.base[field_of(q)], at);
                }
            }
#pragma unroll
            for (int ii = 0; ii < PF; ++ii) {
                const int m = tid + NT * (i0 + ii);
                if (i0 + ii >= ACC || m >= nmodes) break;
                const int c = divn(m, nay, d.lnay);
                const int cidx = m - c * nay;
                const int gx = gx0 + c;
                if (gx >= hxa) continue;
                const uint32_t at = static_cast<uint32_t>(gx * nay + cidx);
                if (cidx == half) {   // the always-zero ky = -Nay/2 slot
                    if constexpr (kPsi) {
                        gstore(P.next[1 + a], at, czero<T>());
                    } else {
                        if (compressible) gstore(P.next[0], at, czero<T>());
                        gstore(P.next[fx_field], at, czero<T>());
                        gstore(P.next[fy_field], at, czero<T>());
                    }
                    continue;
                }
                const int ky_i = cidx < half ? cidx : cidx - nay;
                const T kx = dkx * T(gx);
                const T ky = dky * T(ky_i);
                const T k2 = kx * kx + ky * ky;
                const bool is_k0 = kFp32 && gx == 0 && cidx == 0 && P.next_k0 != nullptr;
                // canonical mode of the noise draw: the kx = 0 line draws the ky > 0
                // member of each pair and conjugates it onto ky < 0
                // (src/fcalculator.cc:525-566)
                const bool conj_draw = (gx == 0 && ky_i < 0);
                const uint32_t mode_id = static_cast<uint32_t>(gx * nay + (conj_draw ? -ky_i : cidx));

                // every load of the mode is issued here, ahead of the noise draws and
                // the flux updates (whose stores the later loads could not pass);
                // the base state is read only when its weight is nonzero
                if constexpr (kPsi) {
                    const int f = 1 + a;
                    const cx<T> cur = pcur[ii][0];
                    const cx<T> bas = pbas[ii][0];
                    cx<T> det = acc_sum(acc_x, m);
                    if (d.psi_det_on[a]) {
                        if (d.psi_lin_on[a]) {   // src/fcalculator.cc:152-187
                            const T mob = T(d.mobility[a]);
                            const cx<T> mu = (T(d.mu_k0[a]) + T(d.mu_k2[a]) * k2 + T(d.mu_k4[a]) * k2 * k2) * cur;
                            const cx<T> fx = mul_mi((mob * kx) * mu);
                            const cx<T> fy = mul_mi((mob * ky) * mu);
                            det = det + mul_mi(kx * fx + ky * fy);
                            if (fx_slot) {
                                fx_slot[at] = fx_slot[at] + fw * fx;
                                fy_slot[at] = fy_slot[at] + fw * fy;
                            }
                        }
                        for (int q = 0; q < d.n_igrad; ++q) {   // src/fcalculator.cc:797-812
                            if (d.igrad_component[q] != a) continue;
                            const cx<T> j = gload(P.cur[d.igrad_direction[q] == 0 ? fx_field : fy_field], at);
                            const cx<T> t = T(-d.igrad_amplitude[q]) * j;
                            det = det + mk(t.re / T(d.rho0), t.im / T(d.rho0));
                        }
                        for (int q = 0; q < d.nsources; ++q) {
                            if (d.sources[q].field == f && d.sources[q].gx == gx && d.sources[q].cidx == cidx)
                                det = det + mk(T(d.sources[q].re), T(d.sources[q].im));
                        }
                    }
                    cx<T> sto = czero<T>();
                    if (d.psi_noise_on[a] && k2 != T(0)) {   // src/fcalculator.cc:568-584
                        // draws A and B (conjugated on the kx = 0, ky < 0 half line);
                        // the counter-based generator gives every SRK3 stage of a
                        // step the same A and B, so each stage draws them itself
                        cx<T> sa, sb;
                        const T sigma = sqrt(T(d.psi_noise_pref[a]) * k2);   // in T: FP32 needs no FP64 square root
                        gauss_quad<T>(d, mode_id, static_cast<uint32_t>(a), sp, sa, sb);   // draws A and B
                        sa = sigma * sa;
                        sb = sp.use_b ? sigma * sb : czero<T>();
                        if (conj_draw) { sa = conj(sa); sb = conj(sb); }
                        sto = sp.use_b ? sa + beta * sb : sa;
                        if (fx_slot && sp.flux_first) {   // draw A: (i k / k^2) sigma xi, times 1/sqrt(dt)
                            const T isd = T(1.0 / d.sqrt_dt);
                            fx_slot[at] = fx_slot[at] + isd * mul_i((kx / k2) * sa);
                            fy_slot[at] = fy_slot[at] + isd * mul_i((ky / k2) * sa);
                        }
                    }
                    if (is_k0) k0_write<T>(P.base_k0, P.cur_k0, P.next_k0, P.next[f] + at, f, dt * det, sp);
                    else gstore(P.next[f], at, wb * bas + wc * (cur + dt * det + sdt * sto));
                } else {
                    const bool pressure = d.mom_det_on && d.pressure_coefficient != 0.0;
                    const cx<T> jx = pcur[ii][1];
                    const cx<T> jy = pcur[ii][2];
                    const cx<T> rho = pcur[ii][0];
                    const cx<T> bx = pbas[ii][1];
                    const cx<T> by = pbas[ii][2];
                    const cx<T> brho = pbas[ii][0];
                    cx<T> dx = acc_sum(acc_x, m), dy = acc_sum(acc_y, m);
                    if (d.mom_det_on) {
                        if (pressure) {   // src/fcalculator.cc:276-300
                            const cx<T> p = T(d.pressure_coefficient) * rho;
                            dx = dx + mul_mi(kx * p);
                            dy = dy + mul_mi(ky * p);
                            if (pi_slot) {
                                pi_slot[at] = pi_slot[at] + fw * p;
                                pi_slot[at + 2 * fe] = pi_slot[at + 2 * fe] + fw * p;
                            }
                        }
                        if (d.incompressible_viscous) {   // src/fcalculator.cc:302-341
                            const T eta = T(d.eta);
                            const T r0 = T(d.rho0);
                            const cx<T> vx = mk(jx.re / r0, jx.im / r0);
                            const cx<T> vy = mk(jy.re / r0, jy.im / r0);
                            const cx<T> pxx = mul_mi(eta * ((T(2) * kx) * vx));
                            const cx<T> pxy = mul_mi(eta * (ky * vx + kx * vy));
                            const cx<T> pyy = mul_mi(eta * ((T(2) * ky) * vy));
                            dx = dx + mul_mi(kx * pxx + ky * pxy);
                            dy = dy + mul_mi(kx * pxy + ky * pyy);
                            if (pi_slot) {
                                pi_slot[at] = pi_slot[at] + fw * pxx;
                                pi_slot[at + fe] = pi_slot[at + fe] + fw * pxy;
                                pi_slot[at + 2 * fe] = pi_slot[at + 2 * fe] + fw * pyy;
                            }
                        }
                        for (int q = 0; q < d.nsources; ++q) {
                            if (d.sources[q].gx != gx || d.sources[q].cidx != cidx) continue;
                            const cx<T> sv = mk(T(d.sources[q].re), T(d.sources[q].im));
                            if (d.sources[q].field == fx_field) dx = dx + sv;
                            if (d.sources[q].field == fy_field) dy = dy + sv;
                        }
                    }
                    cx<T> sx = czero<T>(), sy = czero<T>();
                    if (d.j_noise_on && k2 != T(0)) {   // src/fcalculator.cc:586-700
                        const T ss = T(d.j_sigma), se = T(d.j_sqrt_eta), sz = T(d.j_sqrt_zeta);
                        const T kxd = conj_draw ? T(0) : kx;
                        const T kyd = conj_draw ? -ky : ky;
                        const int ndraw = sp.use_b ? 2 : 1;
                        cx<T> x3a, x3b;   // the xy deviates of draws A and B
                        gauss_quad<T>(d, mode_id, 10u, sp, x3a, x3b);
                        for (int ab = 0; ab < ndraw; ++ab) {
                            cx<T> x1, x2;
                            gauss_quad<T>(d, mode_id, 8u + static_cast<uint32_t>(ab), sp, x1, x2);
                            const cx<T> x3 = ab == 0 ? x3a : x3b;
                            const cx<T> pxx = -ss * (sz * x1 + se * x2);
                            const cx<T> pyy = -ss * (sz * x1 - se * x2);
                            const cx<T> pxy = (-ss * se) * x3;
                            cx<T> vx = mul_mi(kxd * pxx + kyd * pxy);
                            cx<T> vy = mul_mi(kxd * pxy + kyd * pyy);
                            if (conj_draw) { vx = conj(vx); vy = conj(vy); }
                            if (ab == 0 && pi_slot && sp.flux_first) {   // draw A stress, times 1/sqrt(dt)
                                const T isd = T(1.0 / d.sqrt_dt);
                                pi_slot[at] = pi_slot[at] + isd * (conj_draw ? conj(pxx) : pxx);
                                pi_slot[at + fe] = pi_slot[at + fe] + isd * (conj_draw ? conj(pxy) : pxy);
                                pi_slot[at + 2 * fe] = pi_slot[at + 2 * fe] + isd * (conj_draw ? conj(pyy) : pyy);
                            }
                            if (ab == 0) { sx = vx; sy = vy; }
                            else { sx = sx + beta * vx; sy = sy + beta * vy; }
                        }
                    }
                    if (pi_slot && sp.flux_last && !compressible && k2 != T(0)) {
                        // FluxBuffer::add_incompressible_projection_pressure_flux
                        // (src/buffer_flux.cc:196-225) on the step's accumulated stress
                        const cx<T> txx = pi_slot[at], txy = pi_slot[at + fe], tyy = pi_slot[at + 2 * fe];
                        const cx<T> rx = mul_mi(kx * txx + ky * txy);
                        const cx<T> ry = mul_mi(kx * txy + ky * tyy);
                        const cx<T> kr = kx * rx + ky * ry;
                        const cx<T> pk = mul_mi(mk(kr.re / k2, kr.im / k2));
                        pi_slot[at] = txx + pk;
                        pi_slot[at + 2 * fe] = tyy + pk;
                    }
                    if (compressible) {
                        const cx<T> drho = mul_mi(kx * jx + ky * jy);   // src/fcalculator.cc:60-84
                        if (fx_slot) {
                            fx_slot[at] = fx_slot[at] + fw * jx;
                            fy_slot[at] = fy_slot[at] + fw * jy;
                        }
                        if (is_k0) {
                            k0_write<T>(P.base_k0, P.cur_k0, P.next_k0, P.next[0] + at, 0, dt * drho, sp);
                            k0_write<T>(P.base_k0, P.cur_k0, P.next_k0, P.next[fx_field] + at, fx_field, dt * dx, sp);
                            k0_write<T>(P.base_k0, P.cur_k0, P.next_k0, P.next[fy_field] + at, fy_field, dt * dy, sp);
                        } else {
                            gstore(P.next[0], at, wb * brho + wc * (rho + dt * drho));
                            gstore(P.next[fx_field], at, wb * bx + wc * (jx + dt * dx + sdt * sx));
                            gstore(P.next[fy_field], at, wb * by + wc * (jy + dt * dy + sdt * sy));
                        }
                    } else {   // incompressible: projected increment (time_integrator_srk3_incompressible.h:81-107)
                        cx<T> ix = dt * dx + sdt * sx;
                        cx<T> iy = dt * dy + sdt * sy;
                        if (k2 != T(0)) {
                            const cx<T> tr = (-ky) * ix + kx * iy;
                            const cx<T> px = (-ky) * tr;
                            const cx<T> py = kx * tr;
                            ix = mk(px.re / k2, px.im / k2);
                            iy = mk(py.re / k2, py.im / k2);
                        }
                        if (is_k0) {
                            k0_write<T>(P.base_k0, P.cur_k0, P.next_k0, P.next[fx_field] + at, fx_field, ix, sp);
                            k0_write<T>(P.base_k0, P.cur_k0, P.next_k0, P.next[fy_field] + at, fy_field, iy, sp);
                        } else {
                            gstore(P.next[fx_field], at, wb * bx + wc * (jx + ix));
                            gstore(P.next[fy_field], at, wb * by + wc * (jy + iy));
                        }
                    }
                }
            }
        }
    };
    if (psi_group) stage_modes(std::true_type{});
    else stage_modes(std::false_type{});

    // ---- Hermitian projection of the kx = 0 line (src/time_integrator.h:70-103)
    // The block owning gx = 0 re-reads its own writes after a workgroup barrier
    // (which orders the global stores within the workgroup); each (ky, -ky)
    // pair is handled by one thread.
    if (gx0 == 0) {
        __threadfence_block();
        __syncthreads();
        int nf = 0;
        int wf0 = 0, wf1 = 0, wf2 = 0;
        if (psi_group) { wf0 = 1 + a; nf = 1; }
        else if (compressible) { wf0 = 0; wf1 = fx_field; wf2 = fy_field; nf = 3; }
        else { wf0 = fx_field; wf1 = fy_field; nf = 2; }
        for (int q = 0; q < nf; ++q) {
            cx<T>* col = P.next[q == 0 ? wf0 : (q == 1 ? wf1 : wf2)];   // column gx = 0 starts at index 0
            for (int m = tid; m < half; m += NT) {
                if (m == 0) {
                    col[0].im = T(0);
                } else {
                    const cx<T> av = col[m];
                    const cx<T> bv = col[nay - m];
                    const cx<T> p = T(0.5) * (av + conj(bv));
                    col[m] = p;
                    col[nay - m] = conj(p);
                }
            }
        }
    }
}

}  // namespace shhd
