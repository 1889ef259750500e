// measure.h — FP64 on-device accumulators that run after every step, so the
// measures that read a few modes or a per-mode average every step need no
// per-step download of the state or the flux:
//
//   * mode probes: the values of chosen (state field or flux slot, mode)
//     pairs after each step, appended to a ring of FP64 complex rows.  This is
//     what TimeSeriesMeasure's pi_* / flux columns (the k = 0 flux,
//     src/measure_time_series.cc:189-227), YokotaGreenKuboMeasure (Pi_xy at
//     chosen wave vectors, src/measure_yokota_green_kubo.cc:128-193) and a
//     low-mode time correlation read each step.
//   * structure factor: sum over samples of |f_hat(k)|^2 for chosen fields at
//     every active mode, sampled every `nevery` steps, in FP64.
#pragma once

#include "desc.h"

namespace shhd {

constexpr int kMaxProbes = 256;

struct Probe {
    int src;        // 0: state field, 1: flux slot
    int index;      // field or slot
    int off;        // compact offset gx * Nay + c
};

template <typename T>
struct MeasDev {
    // probes
    int nprobe;
    int capacity;               // rows in the ring
    const Probe* probes;
    cx<double>* ring;           // [capacity][nprobe]
    int* rows;                  // rows written since the last read
    const cx<T>* flux;
    // structure factor
    int sk_nf;
    int sk_field[kMaxFields];
    int sk_nevery;
    uint64_t sk_base;           // step count when accumulation started
    double* sk;                 // [sk_nf][field_elems]
    unsigned long long* sk_count;
    // common
    const uint64_t* step;       // steps completed before the current one
    const int* flag;            // zero-density flag: set, the instance took no step
    size_t field_elems;
    unsigned static_mask;       // fields stored once (buffer 0) whatever the rotation
    const cx<double>* k0_static;   // FP32 form: FP64 k = 0 modes of buffer 0, else null
};

// FP64 k = 0 copy of field f after the step (FP32 form), or null
template <typename T>
__device__ __forceinline__ const cx<double>* k0_of(const MeasDev<T>& M, const StagePtrs<T>& P, int f) {
    return ((M.static_mask >> f) & 1u) ? M.k0_static : P.next_k0;
}

// One workgroup per instance: copy the probed values of the state the step
// just wrote (and of its flux) into the next ring row.
template <typename T>
__global__ __launch_bounds__(256) void k_probe(const MeasDev<T>* __restrict__ meas,
                                               const StagePtrs<T>* __restrict__ last) {
    const MeasDev<T>& M = meas[blockIdx.x];
    if (M.nprobe == 0 || instance_failed(M.flag)) return;
    const StagePtrs<T>& P = last[blockIdx.x];
    const int row = *M.rows;
    if (row < M.capacity) {
        for (int p = threadIdx.x; p < M.nprobe; p += blockDim.x) {
            const Probe q = M.probes[p];
            cx<double> v;
            if (q.src == 0) {
                const cx<double>* k0 = k0_of(M, P, q.index);
                if (q.off == 0 && k0 != nullptr) {
                    v = k0[q.index];
                } else {
                    const cx<T> t = P.next[q.index][q.off];
                    v = mk(static_cast<double>(t.re), static_cast<double>(t.im));
                }
            } else {
                const cx<T> t = M.flux[static_cast<size_t>(q.index) * M.field_elems + q.off];
                v = mk(static_cast<double>(t.re), static_cast<double>(t.im));
            }
            M.ring[static_cast<size_t>(row) * M.nprobe + p] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *M.rows = row + 1;
}

// blockIdx.y = instance; the x blocks stride over its sampled fields' modes.
template <typename T>
__global__ __launch_bounds__(256) void k_structure_factor(const MeasDev<T>* __restrict__ meas,
                                                          const StagePtrs<T>* __restrict__ last) {
    const MeasDev<T>& M = meas[blockIdx.y];
    if (M.sk_nf == 0 || instance_failed(M.flag)) return;
    const uint64_t done = *M.step + 1 - M.sk_base;   // steps since accumulation started, this one included
    if (done % static_cast<uint64_t>(M.sk_nevery) != 0) return;
    const StagePtrs<T>& P = last[blockIdx.y];
    const size_t fe = M.field_elems;
    const size_t total = static_cast<size_t>(M.sk_nf) * fe;
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < total;
         i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int q = static_cast<int>(i / fe);
        const size_t off = i - static_cast<size_t>(q) * fe;
        const int f = M.sk_field[q];
        double re, im;
        const cx<double>* k0 = k0_of(M, P, f);
        if (off == 0 && k0 != nullptr) {
            re = k0[f].re;
            im = k0[f].im;
        } else {
            const cx<T> t = P.next[f][off];
            re = static_cast<double>(t.re);
            im = static_cast<double>(t.im);
        }
        M.sk[i] += re * re + im * im;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *M.sk_count += 1ull;
}

}  // namespace shhd
