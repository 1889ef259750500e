// lds_fft.h — batched complex FFT of sequences held in LDS, executed by one
// workgroup of NT threads.
//
// Stockham autosort formulation with radix-2/3/4/8 passes: pass p with radix R
// and current sub-transform length ns turns butterfly j of a sequence of
// length L (j < L/R, k = j mod ns) into
//     v_q = x[j + q L/R] * w_L^{sign q k L/(ns R)},   q < R
//     y[(j/ns) ns R + k + p ns] = sum_q v_q w_R^{sign p q},   p < R
// so the output is in natural order after the last pass.  Each pass reads all
// of a thread's butterfly inputs into registers, synchronises, then writes the
// outputs back in place, so one LDS buffer of nseq*L values suffices.
//
// Twiddles come from a table tw[j] = exp(-2 pi i j / L), j < L, computed in
// FP64 on the host (and rounded for the FP32 form).
#pragma once

#include <utility>

#include "cx.h"

namespace shhd {

constexpr int kMaxPasses = 16;

// LDS slot of FFT element i: one pad slot after every 16 elements.  A 16-byte
// LDS access of a wave is conflict-free when each 8 consecutive lanes hit 8
// distinct 16-byte slots mod 8 (measured: 6 cycles per wave access; 10 for the
// stride-8 writes of a first radix-8 pass under this padding, 48 unpadded).
// A pad every 8 elements would make those conflict-free too, but the larger
// buffers cost resident workgroups where LDS limits them (measured slower).
constexpr int kPadShift = 4;
constexpr int kPad = 1 << kPadShift;
__host__ __device__ __forceinline__ int lp(int i) { return i + (i >> kPadShift); }
// LDS values needed for n padded elements
__host__ __device__ __forceinline__ int lp_size(int n) { return n + (n >> kPadShift) + 1; }

// One Stockham pass: radix R applied to sub-transforms of length ns (a power of
// two: radix-3 passes run last), m = L/R butterflies per sequence, twiddle
// stride tstep = L/(ns R).  inv_m is 1/m for the sequence split of batched
// transforms.
struct FftPass {
    int radix, m, lns, tstep;
    float inv_m;
};

struct FftPlanDev {
    int npass;
    FftPass pass[kMaxPasses];
};

template <typename T, int R>
__device__ __forceinline__ void dft_small(cx<T> (&v)[R], int sign) {
    if constexpr (R == 2) {
        const cx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if constexpr (R == 3) {
        const T s3 = T(0.86602540378443864676);   // sin(2 pi / 3)
        const cx<T> t1 = v[1] + v[2];
        const cx<T> t2 = v[0] - T(0.5) * t1;
        const cx<T> d = v[1] - v[2];
        // sign * i * s3 * d
        const cx<T> r = (sign > 0) ? mul_i(s3 * d) : mul_mi(s3 * d);
        v[0] = v[0] + t1;
        v[1] = t2 + r;
        v[2] = t2 - r;
    } else if constexpr (R == 4) {
        const cx<T> a0 = v[0] + v[2], a1 = v[0] - v[2];
        const cx<T> b0 = v[1] + v[3], b1 = v[1] - v[3];
        const cx<T> r = (sign > 0) ? mul_i(b1) : mul_mi(b1);
        v[0] = a0 + b0;
        v[2] = a0 - b0;
        v[1] = a1 + r;
        v[3] = a1 - r;
    } else if constexpr (R == 8) {
        cx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cx<T> o[4] = {v[1], v[3], v[5], v[7]};
        dft_small<T, 4>(e, sign);
        dft_small<T, 4>(o, sign);
        const T h = T(0.70710678118654752440);
        // w8^p o_p with w8 = exp(sign i pi / 4)
        const cx<T> o1 = (sign > 0) ? mk(h * (o[1].re - o[1].im), h * (o[1].re + o[1].im))
                                    : mk(h * (o[1].re + o[1].im), h * (o[1].im - o[1].re));
        const cx<T> o2 = (sign > 0) ? mul_i(o[2]) : mul_mi(o[2]);
        const cx<T> o3 = (sign > 0) ? mk(-h * (o[3].re + o[3].im), h * (o[3].re - o[3].im))
                                    : mk(h * (o[3].im - o[3].re), -h * (o[3].re + o[3].im));
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[5] = e[1] - o1;
        v[2] = e[2] + o2;
        v[6] = e[2] - o2;
        v[3] = e[3] + o3;
        v[7] = e[3] - o3;
    }
}

// Twiddles w^(q t), q = 1..R-1, from the table tw[j] = exp(-2 pi i j / L), in
// two steps so the table loads can be issued before a pass waits at its
// barrier: load_twiddles fetches w^t, w^2t, w^4t for radix 8 (the other four
// are formed by products: error growth of at most two roundings), w^t, w^2t for
// radix 4 and 3, w^t for 2, conjugated for sign > 0; mul_twiddles applies them
// (t = 0 gives w = 1 exactly).
template <typename T, int R>
struct Twiddles {
    cx<T> w[R == 8 ? 3 : (R == 2 ? 1 : 2)];
};
template <typename T, int R>
__device__ __forceinline__ Twiddles<T, R> load_twiddles(const cx<T>* __restrict__ tw, int t, int sign) {
    auto ld = [&](int e) {
        cx<T> w = gload(tw, static_cast<uint32_t>(e));
        if (sign > 0) w.im = -w.im;
        return w;
    };
    Twiddles<T, R> r;
    if constexpr (R == 8) {
        r.w[0] = ld(t); r.w[1] = ld(2 * t); r.w[2] = ld(4 * t);
    } else if constexpr (R == 2) {
        r.w[0] = ld(t);
    } else {
        r.w[0] = ld(t); r.w[1] = ld(2 * t);
    }
    return r;
}
template <typename T, int R>
__device__ __forceinline__ void mul_twiddles(cx<T> (&v)[R], const Twiddles<T, R>& r) {
    if constexpr (R == 8) {
        const cx<T> w1 = r.w[0], w2 = r.w[1], w4 = r.w[2];
        const cx<T> w3 = w1 * w2, w5 = w1 * w4, w6 = w2 * w4, w7 = w3 * w4;
        v[1] = v[1] * w1; v[2] = v[2] * w2; v[3] = v[3] * w3; v[4] = v[4] * w4;
        v[5] = v[5] * w5; v[6] = v[6] * w6; v[7] = v[7] * w7;
    } else if constexpr (R == 4) {
        v[1] = v[1] * r.w[0]; v[2] = v[2] * r.w[1]; v[3] = v[3] * (r.w[0] * r.w[1]);
    } else if constexpr (R == 3) {
        v[1] = v[1] * r.w[0]; v[2] = v[2] * r.w[1];
    } else {
        v[1] = v[1] * r.w[0];
    }
}

// tid: the lane within the NT lanes that run the transform (a workgroup may
// hold several such groups, each with its own buffer)
template <typename T, int NT, int VPT, int R>
__device__ __forceinline__ void fft_pass(cx<T>* buf, int L, int nseq, const FftPass& ps, const cx<T>* __restrict__ tw,
                                         int sign, int tid) {
    constexpr int B = (VPT + R - 1) / R;
    const int m = ps.m;
    const int nb = nseq * m;
    const int ns_mask = (1 << ps.lns) - 1;
    cx<T> v[B][R];
    int jj[B], base[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int id = tid + NT * b;
        jj[b] = 0;
        base[b] = 0;
        if (id < nb) {
            const int s = (nseq == 1) ? 0 : static_cast<int>((static_cast<float>(id) + 0.5f) * ps.inv_m);
            const int j = id - s * m;
            jj[b] = j;
            base[b] = s * L;
#pragma unroll
            for (int q = 0; q < R; ++q) v[b][q] = buf[lp(s * L + j + q * m)];
        }
    }
    // twiddle loads issued before the barrier (they do not depend on LDS)
    Twiddles<T, R> twv[B];
#pragma unroll
    for (int b = 0; b < B; ++b) twv[b] = load_twiddles<T, R>(tw, (jj[b] & ns_mask) * ps.tstep, sign);
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int id = tid + NT * b;
        if (id < nb) {
            const int j = jj[b];
            const int k = j & ns_mask;
            if (ps.lns > 0) mul_twiddles<T, R>(v[b], twv[b]);
            dft_small<T, R>(v[b], sign);
            const int out = base[b] + ((j >> ps.lns) << ps.lns) * R + k;
#pragma unroll
            for (int p = 0; p < R; ++p) buf[lp(out + (p << ps.lns))] = v[b][p];
        }
    }
    __syncthreads();
}

// In-place FFT of nseq sequences of length L stored back to back in buf.
// Requires nseq * L <= NT * VPT and the plan's radices to multiply to L.
template <typename T, int NT, int VPT>
__device__ __forceinline__ void lds_fft(cx<T>* buf, int L, int nseq, const FftPlanDev& plan,
                                        const cx<T>* __restrict__ tw, int sign,
                                        int tid = static_cast<int>(threadIdx.x)) {
    for (int p = 0; p < plan.npass; ++p) {
        const FftPass& ps = plan.pass[p];
        if (ps.radix == 8) fft_pass<T, NT, VPT, 8>(buf, L, nseq, ps, tw, sign, tid);
        else if (ps.radix == 4) fft_pass<T, NT, VPT, 4>(buf, L, nseq, ps, tw, sign, tid);
        else if (ps.radix == 3) fft_pass<T, NT, VPT, 3>(buf, L, nseq, ps, tw, sign, tid);
        else fft_pass<T, NT, VPT, 2>(buf, L, nseq, ps, tw, sign, tid);
    }
}

// ---------------------------------------------------------------------------
// The same transform with the length fixed at compile time.  Every pass's
// radix, stride and twiddle step are constants, so the sequence split is a
// multiply by a constant, the LDS offsets of a butterfly's R values are
// immediates (lp is linear over a stride that is a multiple of 16), and the
// pass loop and radix dispatch disappear.  Used for the 3*2^k lengths of the
// 3/2-padded grids; every other length takes the runtime plan above.
// ---------------------------------------------------------------------------
constexpr int ct_log2(int v) {
    int l = 0;
    while ((1 << l) < v) ++l;
    return l;
}
// radix of pass P of length L: the host plan's rule (8s, then 4, 2, then 3)
constexpr int ct_radix(int L, int P) {
    int m = L, p = 0;
    while (m % 8 == 0) { if (p == P) return 8; m /= 8; ++p; }
    while (m % 4 == 0) { if (p == P) return 4; m /= 4; ++p; }
    while (m % 2 == 0) { if (p == P) return 2; m /= 2; ++p; }
    while (m % 3 == 0) { if (p == P) return 3; m /= 3; ++p; }
    return 0;
}
constexpr int ct_npass(int L) {
    int n = 0;
    while (ct_radix(L, n) != 0) ++n;
    return n;
}
constexpr int ct_ns(int L, int P) {
    int ns = 1;
    for (int p = 0; p < P; ++p) ns *= ct_radix(L, p);
    return ns;
}

template <typename T, int NT, int VPT, int L, int R, int NS>
__device__ __forceinline__ void fft_pass_ct(cx<T>* buf, int nseq, const cx<T>* __restrict__ tw, int sign, int tid) {
    constexpr int M = L / R;
    constexpr int LNS = ct_log2(NS);
    constexpr int TSTEP = L / (NS * R);
    constexpr int B = (VPT + R - 1) / R;
    const int nb = nseq * M;
    cx<T> v[B][R];
    int jj[B], sl[B];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int id = tid + NT * b;
        jj[b] = 0;
        sl[b] = 0;
        if (id < nb) {
            const int s = id / M;
            const int j = id - s * M;
            jj[b] = j;
            sl[b] = s * L;
            const int a = s * L + j;
            if constexpr (M % kPad == 0) {
                const int a0 = lp(a);
#pragma unroll
                for (int q = 0; q < R; ++q) v[b][q] = buf[a0 + q * (M + M / kPad)];
            } else {
#pragma unroll
                for (int q = 0; q < R; ++q) v[b][q] = buf[lp(a + q * M)];
            }
        }
    }
    // twiddle loads issued before the barrier (they do not depend on LDS)
    Twiddles<T, R> twv[B];
    if constexpr (NS > 1) {
#pragma unroll
        for (int b = 0; b < B; ++b) twv[b] = load_twiddles<T, R>(tw, (jj[b] & (NS - 1)) * TSTEP, sign);
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int id = tid + NT * b;
        if (id < nb) {
            const int j = jj[b];
            const int k = j & (NS - 1);
            if constexpr (NS > 1) mul_twiddles<T, R>(v[b], twv[b]);
            dft_small<T, R>(v[b], sign);
            const int out = sl[b] + ((j >> LNS) << LNS) * R + k;
            if constexpr (NS % kPad == 0) {
                const int o0 = lp(out);
#pragma unroll
                for (int p = 0; p < R; ++p) buf[o0 + p * (NS + NS / kPad)] = v[b][p];
            } else {
#pragma unroll
                for (int p = 0; p < R; ++p) buf[lp(out + p * NS)] = v[b][p];
            }
        }
    }
    __syncthreads();
}

template <typename T, int NT, int VPT, int L, int... P>
__device__ __forceinline__ void lds_fft_ct_passes(cx<T>* buf, int nseq, const cx<T>* __restrict__ tw, int sign,
                                                  int tid, std::integer_sequence<int, P...>) {
    (fft_pass_ct<T, NT, VPT, L, ct_radix(L, P), ct_ns(L, P)>(buf, nseq, tw, sign, tid), ...);
}

template <typename T, int NT, int VPT, int L>
__device__ __forceinline__ void lds_fft_ct(cx<T>* buf, int nseq, const cx<T>* __restrict__ tw, int sign,
                                           int tid = static_cast<int>(threadIdx.x)) {
    // an opaque copy of the lane index: the constant-folded addresses of this
    // call are then formed here, not hoisted out of the caller's loops and
    // held live across them
    asm volatile("" : "+v"(tid));
    __builtin_assume(tid >= 0 && tid < NT);   // the range the copy hides
    lds_fft_ct_passes<T, NT, VPT, L>(buf, nseq, tw, sign, tid, std::make_integer_sequence<int, ct_npass(L)>{});
}

// Fixed-length transform when L is one of the 3*2^k lengths that fit this
// launch shape, else the runtime plan.
template <typename T, int NT, int VPT>
__device__ __forceinline__ void lds_fft_any(cx<T>* buf, int L, int nseq, const FftPlanDev& plan,
                                            const cx<T>* __restrict__ tw, int sign,
                                            int tid = static_cast<int>(threadIdx.x)) {
    constexpr int CAP = NT * VPT;
#define SHHD_FFT_CT(LEN)                                                  \
    if constexpr (VPT == 6 && LEN <= CAP && 8 * LEN >= CAP) {             \
        if (L == LEN) {                                                   \
            lds_fft_ct<T, NT, VPT, LEN>(buf, nseq, tw, sign, tid);        \
            return;                                                       \
        }                                                                 \
    }
    SHHD_FFT_CT(96) SHHD_FFT_CT(192) SHHD_FFT_CT(384) SHHD_FFT_CT(768)
    SHHD_FFT_CT(1536) SHHD_FFT_CT(3072) SHHD_FFT_CT(6144)
#undef SHHD_FFT_CT
    lds_fft<T, NT, VPT>(buf, L, nseq, plan, tw, sign, tid);
}

}  // namespace shhd
