// instances.h — the (lanes, values-per-lane) shapes the kernels are built for.
// Each instance is compiled once, in one of the inst_*.hip translation units;
// the host code only refers to them.
#pragma once

#include "kernels.h"

// X(NT, V): NT lanes per workgroup, V values per lane (6 for 3*2^a sides, 8 for 2^a)
#define SHHD_FOR_NT_V(X) X(64, 6) X(64, 8) X(256, 6) X(256, 8) X(512, 6) X(512, 8) X(1024, 6) X(1024, 8)

namespace shhd {

#define SHHD_DECL_K1(T, nt, v) \
    extern template __global__ void k_inverse_columns<T, nt, v>(const InstDev<T>*, const StagePtrs<T>*, BlockMap);
#define SHHD_DECL_K3(T, nt, v, a) \
    extern template __global__ void k_forward_columns<T, nt, v, a>(const InstDev<T>*, const StagePtrs<T>*, StageParams, BlockMap);
#define SHHD_DECL_K2(T, nt, v, mn) \
    extern template __global__ void k_rows<T, nt, v, mn, 0>(const InstDev<T>*, BlockMap);
#define SHHD_DECL_K2F(T, nt, v, mn, fl) \
    extern template __global__ void k_rows<T, nt, v, mn, fl>(const InstDev<T>*, BlockMap);

// Row kernels built for one row length (3/2-padded grids of 512^2 .. 4096^2):
// X(NT, V, L), the lane shape the row plan picks for rows of L points.
#define SHHD_FOR_ROW_FL(X) X(256, 6, 768) X(256, 6, 1536) X(512, 6, 3072) X(1024, 6, 6144)
#define SHHD_DECL_K2F_ALL(T, nt, v, fl) \
    SHHD_DECL_K2F(T, nt, v, 0, fl) SHHD_DECL_K2F(T, nt, v, 1, fl) SHHD_DECL_K2F(T, nt, v, 2, fl) \
    SHHD_DECL_K2F(T, nt, v, 4, fl) SHHD_DECL_K2F(T, nt, v, 8, fl)
#define SHHD_DECL_K2F_F64(nt, v, fl) SHHD_DECL_K2F_ALL(double, nt, v, fl)
#define SHHD_DECL_K2F_F32(nt, v, fl) SHHD_DECL_K2F_ALL(float, nt, v, fl)

#define SHHD_DECL_KP(T, nt, v) \
    extern template __global__ void k_rows_physical<T, nt, v>(const InstDev<T>*, T*, int);

#define SHHD_DECL_ALL(T, nt, v)                                                              \
    SHHD_DECL_K1(T, nt, v) SHHD_DECL_K3(T, nt, v, 4) SHHD_DECL_K3(T, nt, v, v) SHHD_DECL_KP(T, nt, v) \
    SHHD_DECL_K2(T, nt, v, 0) SHHD_DECL_K2(T, nt, v, 1) SHHD_DECL_K2(T, nt, v, 2)            \
    SHHD_DECL_K2(T, nt, v, 4) SHHD_DECL_K2(T, nt, v, 8)
#define SHHD_DECL_F64(nt, v) SHHD_DECL_ALL(double, nt, v)
#define SHHD_DECL_F32(nt, v) SHHD_DECL_ALL(float, nt, v)
#ifndef SHHD_DEFINE_INSTANCES
SHHD_FOR_NT_V(SHHD_DECL_F64)
SHHD_FOR_NT_V(SHHD_DECL_F32)
SHHD_FOR_ROW_FL(SHHD_DECL_K2F_F64)
SHHD_FOR_ROW_FL(SHHD_DECL_K2F_F32)
#endif

}  // namespace shhd
