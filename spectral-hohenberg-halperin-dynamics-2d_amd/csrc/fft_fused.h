// fft_fused.h — fixed-length Stockham transforms whose first pass takes its
// inputs from registers and whose last pass may leave its outputs in
// registers, so the kernels load a transform's inputs from HBM straight into
// the first butterflies and consume the last butterflies' outputs without an
// LDS round trip.  The passes in between exchange through LDS, alternating
// between two buffers (one barrier per pass) or reusing one (two barriers).
//
// Same Stockham pass as lds_fft.h (butterfly j of a sequence, k = j mod ns:
// inputs x[j + q L/R] twiddled by w_L^{sign q k L/(ns R)}, outputs at
// (j - k) R + k + p ns), with the sub-length ns no longer required to be a
// power of two, so the radix-3 pass may come first.  Two radix orders:
//   R3 last  (4 or 2, 8s, then 3): the last pass's butterfly j (j < L/3) ends
//            with the outputs j, j + L/3, j + 2L/3;
//   R3 first (3, then 8s, 4, 2): the first pass's butterfly j takes the
//            inputs j, j + L/3, j + 2L/3.
// A transform with radix 3 last followed by one with radix 3 first therefore
// hands each lane the same six points (two radix-3 butterflies of 3 points
// when NT * 6 = nseq * L), which is how the row pass goes from the inverse to
// the forward transform through registers.
#pragma once

#include <type_traits>

#include "lds_fft.h"

namespace shhd {

// Radix 3 last puts the 4 or 2 that the radix-8 passes leave of L/3 first,
// so that a radix-8 pass (at ns = L/24) always precedes the radix-3 pass (the
// pair the cross-lane exchange below merges): 3072 = 2 8 8 8 3, 6144 = 4 8 8 8 3.
constexpr int fz_radix(int L, bool r3_first, int P) {
    if (L % 3 != 0) return ct_radix(L, P);
    if (r3_first) return P == 0 ? 3 : ct_radix(L / 3, P - 1);
    int rest = L / 3;   // a power of two for the 3 * 2^k lengths
    while (rest % 8 == 0) rest /= 8;
    if (rest != 1 && rest != 2 && rest != 4) return ct_radix(L, P);
    int p = P;
    if (rest > 1) {
        if (p == 0) return rest;
        --p;
    }
    int m = L / 3 / rest;
    while (m % 8 == 0) {
        if (p == 0) return 8;
        m /= 8;
        --p;
    }
    return p == 0 ? 3 : 0;
}
constexpr int fz_npass(int L, bool r3_first) {
    int n = 0;
    while (fz_radix(L, r3_first, n) != 0) ++n;
    return n;
}
constexpr int fz_ns(int L, bool r3_first, int P) {
    int ns = 1;
    for (int p = 0; p < P; ++p) ns *= fz_radix(L, r3_first, p);
    return ns;
}

// Twiddles of the radix-8 passes, appended to the flat table
// tw[j] = exp(-2 pi i j / L), j < L: for the radix-3-last order, then the
// radix-3-first order, every radix-8 pass P >= 1 of sub-length ns holds 8 ns
// values, the seven twiddles w^(q k L / (8 ns)), q = 1..7, of each butterfly
// index k < ns (copies of the flat table's entries).  A butterfly loads all
// seven instead of forming four of them as products of three.  The values of
// one q lie in a plane over k, so the lanes of a wave (consecutive k) load
// consecutive entries: in FP64 plane q - 1 at (q - 1) ns + k; in FP32 planes
// of pairs, (w^q, w^(q+1)) for odd q at 2 (((q - 1) / 2) ns + k), one 16-byte
// load (the pair of q = 7 is padded).
constexpr int fz_tw8_count(int L, bool r3_first, int P) {   // butterfly indices of the radix-8 passes before P
    int n = 0;
    for (int p = 1; p < P; ++p)
        if (fz_radix(L, r3_first, p) == 8) n += fz_ns(L, r3_first, p);
    return n;
}
constexpr int fz_tw8_at(int L, bool r3_first, int P) {
    return L + 8 * ((r3_first ? fz_tw8_count(L, false, fz_npass(L, false)) : 0) + fz_tw8_count(L, r3_first, P));
}
constexpr int fz_tw_size(int L) { return fz_tw8_at(L, true, fz_npass(L, true)); }

// Butterflies of a radix-R pass over NSEQ sequences of L points on NT lanes:
// lane t holds butterflies id = t + NT b, b < B; id = s M + j (sequence s,
// butterfly j of M = L/R).
template <int NT, int L, int NSEQ, int R>
struct FzMap {
    static constexpr int M = L / R;
    static constexpr int NB = NSEQ * M;
    static constexpr int B = (NB + NT - 1) / NT;
    __device__ __forceinline__ static bool valid(int id) { return NB % NT == 0 || id < NB; }
    // sequence and butterfly of a valid id (one sequence: no division)
    __device__ __forceinline__ static int seq(int id) { return NSEQ == 1 ? 0 : id / M; }
    __device__ __forceinline__ static int bfly(int id) { return NSEQ == 1 ? id : id - (id / M) * M; }
};

// Two LDS buffers used alternately (NBUF = 2) or one buffer (NBUF = 1).
// wr() is the buffer the next pass writes, rd() the one the last pass wrote.
template <typename T, int NBUF>
struct FzBufs {
    cx<T>* b0;
    cx<T>* b1;
    int w;
    __device__ __forceinline__ cx<T>* wr() const { return (NBUF == 2 && w) ? b1 : b0; }
    __device__ __forceinline__ cx<T>* rd() const { return (NBUF == 2 && !w) ? b1 : b0; }
    __device__ __forceinline__ void flip() { w ^= 1; }
    // before a pass writes: with one buffer, every lane must be done reading it
    __device__ __forceinline__ void before_write() const {
        if constexpr (NBUF == 1) __syncthreads();
    }
};

constexpr bool fz_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// LDS layouts of a transform buffer: plain (slot i), padded (lp(i): one slot
// per 16) or padded per group of G consecutive elements (slot i + i / G).
constexpr int kLayPlain = 0, kLayPad = 1, kLayGroup = 2;
template <int LAY, int G>
__device__ __forceinline__ int fz_slot(int i) {
    if constexpr (LAY == kLayPad) return lp(i);
    else if constexpr (LAY == kLayGroup) return i + i / G;
    else return i;
}

// Layout of the buffer pass P writes: plain after a radix-3 first pass (its
// stride-3 writes are conflict-free unpadded and conflict 2-way padded; the
// next pass reads runs of consecutive elements, conflict-free either way) and
// after an inner pass of sub-length 24 (its writes are runs of 24 consecutive
// elements over consecutive lanes: conflict-free unpadded for 16-byte
// elements, at most 2-way for 8-byte ones, where the padded slot would cost
// four integer operations per output); padded per group of ns R elements
// after a pass whose sub-length ns is not a power of two (the radix-8 pass at
// ns = 3 after a radix-3 first pass: butterfly group m writes the run
// [m ns R, (m + 1) ns R), so every output sits at the group's padded base
// plus a constant, where the lp slot of o + p ns costs four integer
// operations per output; its unpadded stride-3 writes conflict 3-way, the
// group-padded ones at most 2-way on a few lanes, and the next pass, whose
// butterfly stride is a multiple of ns R, reads at a lane base plus
// constants); padded after every other pass (stride 8, 4, 2 writes) and after
// the last pass (the callers read the result in the padded layout).
constexpr int fz_layout_after(int L, bool r3_first, int P) {
    if (P == 0 && fz_radix(L, r3_first, 0) == 3) return kLayPlain;
    if (fz_ns(L, r3_first, P) == 24 && P + 1 < fz_npass(L, r3_first)) return kLayPlain;
    const int ns = fz_ns(L, r3_first, P), g = ns * fz_radix(L, r3_first, P);
    if (P + 1 < fz_npass(L, r3_first) && !fz_pow2(ns) && ns % kPad != 0 && g >= kPad &&
        (L / fz_radix(L, r3_first, P + 1)) % g == 0)
        return kLayGroup;
    return kLayPad;
}
// group of the layout pass P writes (ns R of that pass: the next pass's ns)
constexpr int fz_group_after(int L, bool r3_first, int P) { return fz_ns(L, r3_first, P) * fz_radix(L, r3_first, P); }

// Stockham outputs of the butterflies in v to LDS.
template <typename T, int NT, int L, int NSEQ, int R, int NS, int LAY = kLayPad>
__device__ __forceinline__ void fz_write(cx<T>* __restrict__ out, cx<T> (&v)[FzMap<NT, L, NSEQ, R>::B][R], int tid) {
    using Mp = FzMap<NT, L, NSEQ, R>;
    static_assert(LAY != kLayGroup || L % (NS * R) == 0, "whole groups per sequence");
#pragma unroll
    for (int b = 0; b < Mp::B; ++b) {
        const int id = tid + NT * b;
        if (!Mp::valid(id)) continue;
        const int s = Mp::seq(id);
        const int j = Mp::bfly(id);
        const int k = j % NS;
        const int o = s * L + (j - k) * R + k;
        // With NS and R powers of two (and every sequence starting on a pad
        // block), lp(o + p NS) = lp(o) + p NS + (p NS >> kPadShift) for every
        // butterfly: o = (multiple of NS R) + k with k < NS, so no output run
        // crosses a pad slot it does not account for; the per-output offsets
        // are then compile-time constants.  In the group layout the run
        // o + p NS (p < R) stays inside group o / (NS R) for the same reason.
        if constexpr (LAY == kLayGroup) {
            const int o0 = o + o / (NS * R);
#pragma unroll
            for (int p = 0; p < R; ++p) out[o0 + p * NS] = v[b][p];
        } else if constexpr (LAY == kLayPad &&
                             (NS % kPad == 0 || (fz_pow2(NS) && fz_pow2(R) && (NSEQ == 1 || L % kPad == 0)))) {
            const int o0 = lp(o);
#pragma unroll
            for (int p = 0; p < R; ++p) out[o0 + p * NS + ((p * NS) >> kPadShift)] = v[b][p];
        } else {
#pragma unroll
            for (int p = 0; p < R; ++p) out[fz_slot<LAY, NS * R>(o + p * NS)] = v[b][p];
        }
    }
}

// The seven twiddles of butterfly index k of a radix-8 pass whose planes
// start at `pass` (sub-length NS).
template <typename T>
struct Tw8 {
    cx<T> w[7];
};
template <typename T, int NS>
__device__ __forceinline__ Tw8<T> load_tw8(const cx<T>* __restrict__ pass, int k, int sign) {
    Tw8<T> r;
    if constexpr (sizeof(T) == 4) {
        // a uniform base and a 32-bit lane byte offset (the table is far
        // below 4 GiB), as gload does
        typedef float f4 __attribute__((ext_vector_type(4)));
        typedef float f2 __attribute__((ext_vector_type(2)));
        const char* base = reinterpret_cast<const char*>(pass);
        const uint32_t kb = static_cast<uint32_t>(k) * 16u;
#pragma unroll
        for (int h = 0; h < 3; ++h) {
            const f4 v = *(const __attribute__((address_space(1))) f4*)(base + (kb + uint32_t(h * NS * 16)));
            r.w[2 * h] = mk(v.x, v.y);
            r.w[2 * h + 1] = mk(v.z, v.w);
        }
        // the padded pair of q = 7: only its first twiddle is loaded (a
        // 16-byte load whose unused half is overwritten must complete first)
        const f2 v7 = *(const __attribute__((address_space(1))) f2*)(base + (kb + uint32_t(3 * NS * 16)));
        r.w[6] = mk(v7.x, v7.y);
    } else {
#pragma unroll
        for (int q = 0; q < 7; ++q) r.w[q] = gload(pass, static_cast<uint32_t>(q * NS + k));
    }
#pragma unroll
    for (int q = 0; q < 7; ++q)
        if (sign > 0) r.w[q].im = -r.w[q].im;
    return r;
}
template <typename T>
__device__ __forceinline__ void mul_tw8(cx<T> (&v)[8], const Tw8<T>& r) {
#pragma unroll
    for (int q = 1; q < 8; ++q) v[q] = v[q] * r.w[q - 1];
}

// Table twiddles of the lane's butterflies for pass P (sub-length NS) of the
// given radix order, issued before the pass's barrier (they do not depend on
// the LDS data): the appended planes for radix 8, the flat table
// (w^t, w^2t, and w^4t with products for the rest) otherwise.
// (The FP64 builds of 1024 lanes keep the three-load form: their registers
// are already short, and seven twiddles held across the barrier cost more
// than the four products save.)
template <typename T, int NT, int R>
constexpr bool fz_tw8_planes() { return R == 8 && (sizeof(T) == 4 || NT < 1024); }
template <typename T, int NT, int L, int NSEQ, int R, int NS>
struct FzTw {
    using W = std::conditional_t<fz_tw8_planes<T, NT, R>(), Tw8<T>, Twiddles<T, R>>;
    W t[FzMap<NT, L, NSEQ, R>::B];
};
template <typename T, int NT, int L, int NSEQ, bool R3F, int P>
__device__ __forceinline__ auto fz_twiddles(const cx<T>* __restrict__ tw, int sign, int tid) {
    constexpr int R = fz_radix(L, R3F, P);
    constexpr int NS = fz_ns(L, R3F, P);
    using Mp = FzMap<NT, L, NSEQ, R>;
    constexpr int TSTEP = L / (NS * R);
    FzTw<T, NT, L, NSEQ, R, NS> r;
#pragma unroll
    for (int b = 0; b < Mp::B; ++b) {
        const int id = tid + NT * b;
        const int j = Mp::bfly(id);
        const int k = Mp::valid(id) ? j % NS : 0;
        if constexpr (fz_tw8_planes<T, NT, R>()) r.t[b] = load_tw8<T, NS>(tw + fz_tw8_at(L, R3F, P), k, sign);
        else r.t[b] = load_twiddles<T, R>(tw, k * TSTEP, sign);
    }
    return r;
}

// Butterfly inputs from LDS, twiddled (twiddles loaded beforehand),
// transformed in registers.
// LAY: the layout the previous pass wrote (kLayGroup: groups of NS, this
// pass's sub-length, which divides the butterfly stride M).
template <typename T, int NT, int L, int NSEQ, int R, int NS, int LAY = kLayPad>
__device__ __forceinline__ void fz_read(const cx<T>* __restrict__ in, cx<T> (&v)[FzMap<NT, L, NSEQ, R>::B][R],
                                        const FzTw<T, NT, L, NSEQ, R, NS>& twv, int sign, int tid) {
    using Mp = FzMap<NT, L, NSEQ, R>;
    static_assert(LAY != kLayGroup || (Mp::M % NS == 0 && L % NS == 0), "group layout: whole groups per stride");
#pragma unroll
    for (int b = 0; b < Mp::B; ++b) {
        const int id = tid + NT * b;
        if (!Mp::valid(id)) continue;
        const int s = Mp::seq(id);
        const int j = Mp::bfly(id);
        const int a = s * L + j;
        if constexpr (LAY == kLayGroup) {
            const int a0 = a + a / NS;
#pragma unroll
            for (int q = 0; q < R; ++q) v[b][q] = in[a0 + q * (Mp::M + Mp::M / NS)];
        } else if constexpr (LAY == kLayPad && Mp::M % kPad == 0) {
            const int a0 = lp(a);
#pragma unroll
            for (int q = 0; q < R; ++q) v[b][q] = in[a0 + q * (Mp::M + Mp::M / kPad)];
        } else {
#pragma unroll
            for (int q = 0; q < R; ++q) v[b][q] = in[fz_slot<LAY, NS>(a + q * Mp::M)];
        }
    }
#pragma unroll
    for (int b = 0; b < Mp::B; ++b) {
        const int id = tid + NT * b;
        if (!Mp::valid(id)) continue;
        if constexpr (NS > 1 && fz_tw8_planes<T, NT, R>()) mul_tw8<T>(v[b], twv.t[b]);
        else if constexpr (NS > 1) mul_twiddles<T, R>(v[b], twv.t[b]);
        dft_small<T, R>(v[b], sign);
    }
}

// First pass (ns = 1, no twiddles) on inputs the caller placed in v:
// v[b][q] = x_s[j + q L/R] for butterfly id = tid + NT b = s (L/R) + j.
template <typename T, int NT, int L, int NSEQ, bool R3F, int NBUF>
__device__ __forceinline__ void fz_first(FzBufs<T, NBUF>& bf,
                                         cx<T> (&v)[FzMap<NT, L, NSEQ, fz_radix(L, R3F, 0)>::B][fz_radix(L, R3F, 0)],
                                         int sign, int tid) {
    constexpr int R = fz_radix(L, R3F, 0);
    using Mp = FzMap<NT, L, NSEQ, R>;
    // an opaque copy of the lane index: the constant-folded LDS offsets of
    // each call are formed here, not hoisted out of the caller's loops
    asm volatile("" : "+v"(tid));
    __builtin_assume(tid >= 0 && tid < NT);   // the range the copy hides
#pragma unroll
    for (int b = 0; b < Mp::B; ++b)
        if (Mp::valid(tid + NT * b)) dft_small<T, R>(v[b], sign);
    bf.before_write();
    fz_write<T, NT, L, NSEQ, R, 1, fz_layout_after(L, R3F, 0)>(bf.wr(), v, tid);
    bf.flip();
}

// Pass P (0 < P) from LDS to LDS.
template <typename T, int NT, int L, int NSEQ, bool R3F, int NBUF, int P>
__device__ __forceinline__ void fz_pass(FzBufs<T, NBUF>& bf, const cx<T>* __restrict__ tw, int sign, int tid) {
    constexpr int R = fz_radix(L, R3F, P);
    constexpr int NS = fz_ns(L, R3F, P);
    cx<T> v[FzMap<NT, L, NSEQ, R>::B][R];
    const auto twv = fz_twiddles<T, NT, L, NSEQ, R3F, P>(tw, sign, tid);
    __syncthreads();
    fz_read<T, NT, L, NSEQ, R, NS, fz_layout_after(L, R3F, P - 1)>(bf.rd(), v, twv, sign, tid);
    bf.before_write();
    fz_write<T, NT, L, NSEQ, R, NS, fz_layout_after(L, R3F, P)>(bf.wr(), v, tid);
    bf.flip();
}

template <typename T, int NT, int L, int NSEQ, bool R3F, int NBUF, int... P>
__device__ __forceinline__ void fz_passes(FzBufs<T, NBUF>& bf, const cx<T>* __restrict__ tw, int sign, int tid,
                                          std::integer_sequence<int, P...>) {
    (fz_pass<T, NT, L, NSEQ, R3F, NBUF, P + 1>(bf, tw, sign, tid), ...);
}

// Passes 1 .. npass-1, every one through LDS: the transform ends in LDS
// (bf.rd(), natural order, after a barrier the caller issues).
template <typename T, int NT, int L, int NSEQ, bool R3F, int NBUF>
__device__ __forceinline__ void fz_rest_to_lds(FzBufs<T, NBUF>& bf, const cx<T>* __restrict__ tw, int sign, int tid) {
    asm volatile("" : "+v"(tid));
    __builtin_assume(tid >= 0 && tid < NT);   // the range the copy hides
    fz_passes<T, NT, L, NSEQ, R3F, NBUF>(bf, tw, sign, tid, std::make_integer_sequence<int, fz_npass(L, R3F) - 1>{});
}

// Passes 1 .. npass-1 with the last one left in registers: v[b][p] = y_s[j + p L/R]
// for butterfly id = tid + NT b = s (L/R) + j of the last (radix R) pass.
template <typename T, int NT, int L, int NSEQ, bool R3F, int NBUF>
__device__ __forceinline__ void fz_rest_to_regs(
    FzBufs<T, NBUF>& bf,
    cx<T> (&v)[FzMap<NT, L, NSEQ, fz_radix(L, R3F, fz_npass(L, R3F) - 1)>::B][fz_radix(L, R3F, fz_npass(L, R3F) - 1)],
    const cx<T>* __restrict__ tw, int sign, int tid) {
    constexpr int NP = fz_npass(L, R3F);
    asm volatile("" : "+v"(tid));
    __builtin_assume(tid >= 0 && tid < NT);   // the range the copy hides
    fz_passes<T, NT, L, NSEQ, R3F, NBUF>(bf, tw, sign, tid, std::make_integer_sequence<int, NP - 2>{});
    constexpr int RL = fz_radix(L, R3F, NP - 1), NSL = fz_ns(L, R3F, NP - 1);
    const auto twv = fz_twiddles<T, NT, L, NSEQ, R3F, NP - 1>(tw, sign, tid);
    __syncthreads();
    fz_read<T, NT, L, NSEQ, RL, NSL, fz_layout_after(L, R3F, NP - 2)>(bf.rd(), v, twv, sign, tid);
}

// ---------------------------------------------------------------------------
// Wave-local radix-8 / radix-3 pair.  For L = 3M (M a power of two) with the
// radix-8 pass at ns = L/24 right before the radix-3 pass (radix 3 last), or
// right after it at ns = 3 (radix 3 first), the two passes exchange data only
// inside groups of 3 radix-8 and 8 radix-3 butterflies that share one index
// k < K = L/24.  Lanes are arranged so that a group sits in the four 16-lane
// rows of one wave at the same position kk: group g = 16 (lane / 64) + kk,
// sequence s = g / K, k = g mod K, row r = (lane / 16) mod 4.  Row r < 3 does
// the radix-8 butterfly K r + k (radix 3 last) or 3k + r (radix 3 first);
// every row does the radix-3 butterflies j3 = k + K (r + 4h), h = 0, 1.  The
// exchange between the two passes is then a 4 x 4 transpose of (row, value
// index), done with v_permlane32_swap and v_permlane16_swap (no LDS traffic,
// no barrier).  NT lanes hold NT / 4 groups = NSEQ K, i.e. NT * 6 = NSEQ L.
// ---------------------------------------------------------------------------
template <int L>
struct FzGrp {
    static constexpr int K = L / 24;
    int s, k, r;
    __device__ __forceinline__ explicit FzGrp(int tid) {
        const int g = ((tid >> 6) << 4) | (tid & 15);
        s = g / K;
        k = g - s * K;
        r = (tid >> 4) & 3;
    }
    // radix-3 butterfly of half h
    __device__ __forceinline__ int j3(int h) const { return k + K * (r + 4 * h); }
};

// the pair can replace an LDS exchange for this length and order
template <int L, bool R3F>
constexpr bool fz_pair_ok() {
    if (L % 24 != 0 || fz_npass(L, R3F) < 3) return false;
    if (R3F) return fz_radix(L, true, 0) == 3 && fz_radix(L, true, 1) == 8;
    constexpr int NP = fz_npass(L, false);
    return fz_radix(L, false, NP - 1) == 3 && fz_radix(L, false, NP - 2) == 8 && fz_ns(L, false, NP - 2) == L / 24;
}

// 4 x 4 transpose of (16-lane row, slot) on 32-bit words: afterwards slot q of
// row r holds what slot r of row q held.  permlane32_swap(a, b) exchanges the
// upper half of a's lanes with the lower half of b's (register bit 1 <-> row
// bit 1), permlane16_swap(a, b) the odd rows of a with the even rows of b
// (register bit 0 <-> row bit 0).
__device__ __forceinline__ void fz_xpose4(uint32_t (&x)[4]) {
    const auto p = __builtin_amdgcn_permlane32_swap(x[0], x[2], false, false);
    const auto q = __builtin_amdgcn_permlane32_swap(x[1], x[3], false, false);
    const auto u = __builtin_amdgcn_permlane16_swap(p[0], q[0], false, false);
    const auto v = __builtin_amdgcn_permlane16_swap(p[1], q[1], false, false);
    x[0] = u[0];
    x[1] = u[1];
    x[2] = v[0];
    x[3] = v[1];
}
template <typename T>
__device__ __forceinline__ void fz_xpose4(cx<T> (&v)[4]) {
    constexpr int W = static_cast<int>(sizeof(T) / 4);   // words per real
    uint32_t x[2 * W][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if constexpr (W == 2) {
            const uint64_t re = __builtin_bit_cast(uint64_t, v[q].re), im = __builtin_bit_cast(uint64_t, v[q].im);
            x[0][q] = static_cast<uint32_t>(re);
            x[1][q] = static_cast<uint32_t>(re >> 32);
            x[2][q] = static_cast<uint32_t>(im);
            x[3][q] = static_cast<uint32_t>(im >> 32);
        } else {
            x[0][q] = __builtin_bit_cast(uint32_t, v[q].re);
            x[1][q] = __builtin_bit_cast(uint32_t, v[q].im);
        }
    }
#pragma unroll
    for (int w = 0; w < 2 * W; ++w) fz_xpose4(x[w]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if constexpr (W == 2) {
            v[q].re = __builtin_bit_cast(double, (static_cast<uint64_t>(x[1][q]) << 32) | x[0][q]);
            v[q].im = __builtin_bit_cast(double, (static_cast<uint64_t>(x[3][q]) << 32) | x[2][q]);
        } else {
            v[q].re = __builtin_bit_cast(float, x[0][q]);
            v[q].im = __builtin_bit_cast(float, x[1][q]);
        }
    }
}

// Radix 3 last: passes 1 .. npass-3 through LDS, the radix-8 pass at
// ns = L/24 read from LDS by the group rows, the transpose, and the radix-3
// pass in registers: w[h][p] = y_s[j3(h) + p L/3].
template <typename T, int NT, int L, int NSEQ, int NBUF>
__device__ __forceinline__ void fz_rest_to_regs_pair(FzBufs<T, NBUF>& bf, cx<T> (&w)[2][3],
                                                     const cx<T>* __restrict__ tw, int sign, int tid) {
    static_assert(fz_pair_ok<L, false>() && NT * 6 == NSEQ * L, "radix-8 / radix-3 pair shape");
    constexpr int NP = fz_npass(L, false);
    constexpr int K = L / 24;
    constexpr int M8 = L / 8;                       // radix-8 butterflies per sequence
    constexpr int LAY = fz_layout_after(L, false, NP - 3);
    static_assert(LAY != kLayGroup, "radix 3 last: power-of-two sub-lengths before the pair");
    asm volatile("" : "+v"(tid));
    __builtin_assume(tid >= 0 && tid < NT);   // the range the copy hides
    fz_passes<T, NT, L, NSEQ, false, NBUF>(bf, tw, sign, tid, std::make_integer_sequence<int, NP - 3>{});
    const FzGrp<L> G(tid);
    const bool act = G.r < 3;
    using W8 = std::conditional_t<fz_tw8_planes<T, NT, 8>(), Tw8<T>, Twiddles<T, 8>>;
    W8 t8;
    if constexpr (fz_tw8_planes<T, NT, 8>()) t8 = load_tw8<T, K>(tw + fz_tw8_at(L, false, NP - 2), G.k, sign);
    else t8 = load_twiddles<T, 8>(tw, G.k * (L / (K * 8)), sign);
    const Twiddles<T, 3> t30 = load_twiddles<T, 3>(tw, G.j3(0), sign);
    const Twiddles<T, 3> t31 = load_twiddles<T, 3>(tw, G.j3(1), sign);
    __syncthreads();
    cx<T> v[8];
    if (act) {
        const cx<T>* __restrict__ in = bf.rd();
        const int a = G.s * L + K * G.r + G.k;
        if constexpr (LAY == kLayPad && M8 % kPad == 0) {
            const int a0 = lp(a);
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = in[a0 + q * (M8 + M8 / kPad)];
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = in[fz_slot<LAY, 1>(a + q * M8)];
        }
        if constexpr (fz_tw8_planes<T, NT, 8>()) mul_tw8<T>(v, t8);
        else mul_twiddles<T, 8>(v, t8);
        dft_small<T, 8>(v, sign);
    } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = czero<T>();
    }
    // row a holds X_a[p], p = 0..7; after the transposes row r holds X_a[r + 4h] in slot a
    cx<T> lo[4] = {v[0], v[1], v[2], v[3]};
    cx<T> hi[4] = {v[4], v[5], v[6], v[7]};
    fz_xpose4<T>(lo);
    fz_xpose4<T>(hi);
    cx<T> x0[3] = {lo[0], lo[1], lo[2]};
    cx<T> x1[3] = {hi[0], hi[1], hi[2]};
    mul_twiddles<T, 3>(x0, t30);
    mul_twiddles<T, 3>(x1, t31);
    dft_small<T, 3>(x0, sign);
    dft_small<T, 3>(x1, sign);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        w[0][p] = x0[p];
        w[1][p] = x1[p];
    }
}

// Radix 3 first: the radix-3 pass on the lane's inputs v[h][p] =
// x_s[j3(h) + p L/3], the transpose, and the radix-8 pass at ns = 3 in
// registers, written to LDS; the remaining passes follow with
// fz_rest_to_lds_after_pair.
template <typename T, int NT, int L, int NSEQ, int NBUF>
__device__ __forceinline__ void fz_first_pair(FzBufs<T, NBUF>& bf, cx<T> (&v)[2][3], const cx<T>* __restrict__ tw,
                                              int sign, int tid) {
    static_assert(fz_pair_ok<L, true>() && NT * 6 == NSEQ * L, "radix-3 / radix-8 pair shape");
    asm volatile("" : "+v"(tid));
    __builtin_assume(tid >= 0 && tid < NT);   // the range the copy hides
    const FzGrp<L> G(tid);
    const bool act = G.r < 3;
    const int k8 = act ? G.r : 0;   // twiddle index of radix-8 butterfly 3k + r (ns = 3)
    using W8 = std::conditional_t<fz_tw8_planes<T, NT, 8>(), Tw8<T>, Twiddles<T, 8>>;
    W8 t8;
    if constexpr (fz_tw8_planes<T, NT, 8>()) t8 = load_tw8<T, 3>(tw + fz_tw8_at(L, true, 1), k8, sign);
    else t8 = load_twiddles<T, 8>(tw, k8 * (L / 24), sign);
    dft_small<T, 3>(v[0], sign);
    dft_small<T, 3>(v[1], sign);
    // row r holds Z_{r + 4h}[p3] in slot p3; after the transposes row p3 holds
    // Z_{q}[p3] for q = r + 4h in slot r of half h
    cx<T> lo[4] = {v[0][0], v[0][1], v[0][2], czero<T>()};
    cx<T> hi[4] = {v[1][0], v[1][1], v[1][2], czero<T>()};
    fz_xpose4<T>(lo);
    fz_xpose4<T>(hi);
    cx<T> x[8] = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    if (act) {
        if constexpr (fz_tw8_planes<T, NT, 8>()) mul_tw8<T>(x, t8);
        else mul_twiddles<T, 8>(x, t8);
        dft_small<T, 8>(x, sign);
    }
    bf.before_write();
    if (act) {
        // Stockham output of butterfly j8 = 3k + r at ns = 3: 24 k + r + 3 p
        // (group layout: the run of group s L / 24 + k at its padded base)
        static_assert(fz_layout_after(L, true, 1) == kLayGroup && fz_group_after(L, true, 1) == 24,
                      "the ns = 3 pass writes groups of 24");
        cx<T>* __restrict__ out = bf.wr();
        const int o = G.s * L + 24 * G.k + G.r;
        const int o0 = o + o / 24;
#pragma unroll
        for (int p = 0; p < 8; ++p) out[o0 + 3 * p] = x[p];
    }
    bf.flip();
}

// Passes 2 .. npass-1 through LDS (after fz_first_pair)
template <typename T, int NT, int L, int NSEQ, int NBUF, int... P>
__device__ __forceinline__ void fz_passes_from2(FzBufs<T, NBUF>& bf, const cx<T>* __restrict__ tw, int sign, int tid,
                                                std::integer_sequence<int, P...>) {
    (fz_pass<T, NT, L, NSEQ, true, NBUF, P + 2>(bf, tw, sign, tid), ...);
}
template <typename T, int NT, int L, int NSEQ, int NBUF>
__device__ __forceinline__ void fz_rest_to_lds_after_pair(FzBufs<T, NBUF>& bf, const cx<T>* __restrict__ tw, int sign,
                                                          int tid) {
    asm volatile("" : "+v"(tid));
    __builtin_assume(tid >= 0 && tid < NT);   // the range the copy hides
    fz_passes_from2<T, NT, L, NSEQ, NBUF>(bf, tw, sign, tid, std::make_integer_sequence<int, fz_npass(L, true) - 2>{});
}

}  // namespace shhd
