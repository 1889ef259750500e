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
//   R3 last  (8s, 4, 2, then 3): the last pass's butterfly j (j < L/3) ends
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

constexpr int fz_radix(int L, bool r3_first, int P) {
    if (r3_first && L % 3 == 0) return P == 0 ? 3 : ct_radix(L / 3, P - 1);
    return ct_radix(L, P);
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

// LDS slot of element i of a buffer in the padded (lp) or the plain layout
template <bool PAD>
__device__ __forceinline__ int fz_slot(int i) { return PAD ? lp(i) : i; }

// Layout of the buffer pass P writes: plain after a radix-3 first pass (its
// stride-3 writes are conflict-free unpadded and conflict 2-way padded; the
// next pass reads runs of consecutive elements, conflict-free either way),
// padded after every other pass (stride 8, 4, 2 writes).
constexpr bool fz_pad_after(int L, bool r3_first, int P) { return !(P == 0 && fz_radix(L, r3_first, 0) == 3); }

// Stockham outputs of the butterflies in v to LDS.
template <typename T, int NT, int L, int NSEQ, int R, int NS, bool PAD = true>
__device__ __forceinline__ void fz_write(cx<T>* __restrict__ out, cx<T> (&v)[FzMap<NT, L, NSEQ, R>::B][R], int tid) {
    using Mp = FzMap<NT, L, NSEQ, R>;
#pragma unroll
    for (int b = 0; b < Mp::B; ++b) {
        const int id = tid + NT * b;
        if (!Mp::valid(id)) continue;
        const int s = Mp::seq(id);
        const int j = Mp::bfly(id);
        const int k = j % NS;
        const int o = s * L + (j - k) * R + k;
        if constexpr (PAD && NS % kPad == 0) {
            const int o0 = lp(o);
#pragma unroll
            for (int p = 0; p < R; ++p) out[o0 + p * (NS + NS / kPad)] = v[b][p];
        } else {
#pragma unroll
            for (int p = 0; p < R; ++p) out[fz_slot<PAD>(o + p * NS)] = v[b][p];
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
        typedef float f4 __attribute__((ext_vector_type(4)));
        const __attribute__((address_space(1))) f4* p = (const __attribute__((address_space(1))) f4*)(pass);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const f4 v = p[h * NS + k];
            r.w[2 * h] = mk(v.x, v.y);
            if (h < 3) r.w[2 * h + 1] = mk(v.z, v.w);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 7; ++q) r.w[q] = gload(pass + q * NS + k);
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
template <typename T, int NT, int L, int NSEQ, int R, int NS, bool PAD = true>
__device__ __forceinline__ void fz_read(const cx<T>* __restrict__ in, cx<T> (&v)[FzMap<NT, L, NSEQ, R>::B][R],
                                        const FzTw<T, NT, L, NSEQ, R, NS>& twv, int sign, int tid) {
    using Mp = FzMap<NT, L, NSEQ, R>;
#pragma unroll
    for (int b = 0; b < Mp::B; ++b) {
        const int id = tid + NT * b;
        if (!Mp::valid(id)) continue;
        const int s = Mp::seq(id);
        const int j = Mp::bfly(id);
        const int a = s * L + j;
        if constexpr (PAD && Mp::M % kPad == 0) {
            const int a0 = lp(a);
#pragma unroll
            for (int q = 0; q < R; ++q) v[b][q] = in[a0 + q * (Mp::M + Mp::M / kPad)];
        } else {
#pragma unroll
            for (int q = 0; q < R; ++q) v[b][q] = in[fz_slot<PAD>(a + q * Mp::M)];
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
#pragma unroll
    for (int b = 0; b < Mp::B; ++b)
        if (Mp::valid(tid + NT * b)) dft_small<T, R>(v[b], sign);
    bf.before_write();
    fz_write<T, NT, L, NSEQ, R, 1, fz_pad_after(L, R3F, 0)>(bf.wr(), v, tid);
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
    fz_read<T, NT, L, NSEQ, R, NS, fz_pad_after(L, R3F, P - 1)>(bf.rd(), v, twv, sign, tid);
    bf.before_write();
    fz_write<T, NT, L, NSEQ, R, NS, fz_pad_after(L, R3F, P)>(bf.wr(), v, tid);
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
    fz_passes<T, NT, L, NSEQ, R3F, NBUF>(bf, tw, sign, tid, std::make_integer_sequence<int, NP - 2>{});
    constexpr int RL = fz_radix(L, R3F, NP - 1), NSL = fz_ns(L, R3F, NP - 1);
    const auto twv = fz_twiddles<T, NT, L, NSEQ, R3F, NP - 1>(tw, sign, tid);
    __syncthreads();
    fz_read<T, NT, L, NSEQ, RL, NSL, fz_pad_after(L, R3F, NP - 2)>(bf.rd(), v, twv, sign, tid);
}

}  // namespace shhd
