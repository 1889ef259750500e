// instances.h — the (lanes, values-per-lane) shapes the kernels are built for.
// Each instance is compiled once, in one of the inst_*.hip translation units;
// the host code only refers to them.
#pragma once

#include "kernels.h"

// X(NT, V): NT lanes per workgroup, V values per lane (6 for 3*2^a sides, 8 for 2^a)
#define SHHD_FOR_NT_V(X) X(64, 6) X(64, 8) X(256, 6) X(256, 8) X(512, 6) X(512, 8) X(1024, 6) X(1024, 8)

// Row kernels built for one row length (3/2-padded grids of 512^2 .. 4096^2):
// X(NT, V, L), the lane shape the row plan picks for rows of L points.
#define SHHD_FOR_ROW_FL(X) X(256, 6, 768) X(256, 6, 1536) X(512, 6, 3072) X(1024, 6, 6144)

// Column kernels with fused transforms: X(NT, V, L), C = NT * V / L columns
// of L points per workgroup (3/2-padded sides 64 .. 4096).
#define SHHD_FOR_COL_FZ(X)                                                                             \
    X(256, 6, 96) X(64, 6, 192) X(256, 6, 192) X(64, 6, 384) X(256, 6, 384) X(256, 6, 768)           \
    X(512, 6, 768) X(256, 6, 1536) X(512, 6, 1536) X(512, 6, 3072) X(1024, 6, 3072) X(1024, 6, 6144)

// Forward column kernels whose four wave groups of 64 lanes split the products
// of one column tile (small columns, where a launch has few workgroups): the
// lengths with a one-wave fused build, X(NT, V, L) with NT = 4 * 64 lanes,
// C = 64 * V / L columns per workgroup.
#define SHHD_FOR_COL_FZ_WS(X) X(256, 6, 192) X(256, 6, 384)
// Staged any-length row kernels whose four wave groups of 64 lanes split the
// transforms of one tile of rows (short rows, few workgroups): X(NT, V).
#define SHHD_FOR_ROW_WS(X) X(256, 6)

// order-parameter counts the row kernel is built for (fields held per lane)
#define SHHD_FOR_MAXN(X, ...) X(__VA_ARGS__, 0) X(__VA_ARGS__, 1) X(__VA_ARGS__, 2) X(__VA_ARGS__, 4) X(__VA_ARGS__, 8)
// fixed field layouts of the fused row kernel, (layout, order parameters):
// compressible N = 0, 1, 2; incompressible N = 0; quiescent N = 1
#define SHHD_FOR_LAYOUT(X, ...) X(__VA_ARGS__, 1, 0) X(__VA_ARGS__, 1, 1) X(__VA_ARGS__, 1, 2) X(__VA_ARGS__, 2, 0) \
    X(__VA_ARGS__, 3, 1)

namespace shhd {

// K1<T, NT, V, FZL, NBUF>, K3<T, NT, V, ACC, FZL, NBUF, WS>, K2<T, NT, V, MAXN, FL, NBUF, LAYOUT>
#define SHHD_K1_SIG(T) (const InstDev<T>*, const StagePtrs<T>*, BlockMap)
#define SHHD_K3_SIG(T) (const InstDev<T>*, const StagePtrs<T>*, StageParams, BlockMap)
#define SHHD_K2_SIG(T) (const InstDev<T>*, BlockMap)
#define SHHD_KP_SIG(T) (const InstDev<T>*, T*, int)

// EXT is `extern template` for the declarations, `template` in the unit that builds them
#define SHHD_I_K1(EXT, T, nt, v, fz, nb) EXT __global__ void k_inverse_columns<T, nt, v, fz, nb> SHHD_K1_SIG(T);
#define SHHD_I_K3(EXT, T, nt, v, a, fz, nb) EXT __global__ void k_forward_columns<T, nt, v, a, fz, nb, 1> SHHD_K3_SIG(T);
#define SHHD_I_K3W(EXT, T, nt, v, len)                                                                   \
    EXT __global__ void k_forward_columns<T, nt, v, (2 * (v) / 3), len, 1, 4> SHHD_K3_SIG(T);             \
    EXT __global__ void k_forward_columns<T, nt, v, (2 * (v) / 3), len, 2, 4> SHHD_K3_SIG(T);
#define SHHD_I_K2(EXT, T, nt, v, fl, nb, mn) EXT __global__ void k_rows<T, nt, v, mn, fl, nb, 0, 1> SHHD_K2_SIG(T);
#define SHHD_I_K2L(EXT, T, nt, v, fl, nb, lay, mn) EXT __global__ void k_rows<T, nt, v, mn, fl, nb, lay, 1> SHHD_K2_SIG(T);
#define SHHD_I_K2W(EXT, T, nt, v, mn) EXT __global__ void k_rows<T, nt, v, mn, 0, 0, 0, 4> SHHD_K2_SIG(T);
#define SHHD_I_KP(EXT, T, nt, v) EXT __global__ void k_rows_physical<T, nt, v> SHHD_KP_SIG(T);

// staged column and row kernels of every lane shape, any length
#define SHHD_I_SHAPE(EXT, T, nt, v)                                                              \
    SHHD_I_K1(EXT, T, nt, v, 0, 1) SHHD_I_K3(EXT, T, nt, v, 4, 0, 1) SHHD_I_K3(EXT, T, nt, v, v, 0, 1) \
    SHHD_I_KP(EXT, T, nt, v) SHHD_FOR_MAXN(SHHD_I_K2, EXT, T, nt, v, 0, 0)
// fixed-length row kernels: staged, fused with one and with two LDS buffers
// (any field order, and the fixed layouts)
#define SHHD_I_ROWFL(EXT, T, nt, v, fl)                                                              \
    SHHD_FOR_MAXN(SHHD_I_K2, EXT, T, nt, v, fl, 0) SHHD_FOR_MAXN(SHHD_I_K2, EXT, T, nt, v, fl, 1)  \
    SHHD_FOR_MAXN(SHHD_I_K2, EXT, T, nt, v, fl, 2) SHHD_FOR_LAYOUT(SHHD_I_K2L, EXT, T, nt, v, fl, 1) \
    SHHD_FOR_LAYOUT(SHHD_I_K2L, EXT, T, nt, v, fl, 2)
// fused column kernels, one and two LDS buffers (K3: two modes per radix-3 butterfly)
#define SHHD_I_COLFZ(EXT, T, nt, v, len)                                                               \
    SHHD_I_K1(EXT, T, nt, v, len, 1) SHHD_I_K1(EXT, T, nt, v, len, 2)                                 \
    SHHD_I_K3(EXT, T, nt, v, (2 * (v) / 3), len, 1) SHHD_I_K3(EXT, T, nt, v, (2 * (v) / 3), len, 2)

#define SHHD_EXTERN extern template
#define SHHD_D_SHAPE_F64(nt, v) SHHD_I_SHAPE(SHHD_EXTERN, double, nt, v)
#define SHHD_D_SHAPE_F32(nt, v) SHHD_I_SHAPE(SHHD_EXTERN, float, nt, v)
#define SHHD_D_ROWFL_F64(nt, v, fl) SHHD_I_ROWFL(SHHD_EXTERN, double, nt, v, fl)
#define SHHD_D_ROWFL_F32(nt, v, fl) SHHD_I_ROWFL(SHHD_EXTERN, float, nt, v, fl)
#define SHHD_D_COLFZ_F64(nt, v, len) SHHD_I_COLFZ(SHHD_EXTERN, double, nt, v, len)
#define SHHD_D_COLFZ_F32(nt, v, len) SHHD_I_COLFZ(SHHD_EXTERN, float, nt, v, len)
#define SHHD_D_COLWS_F64(nt, v, len) SHHD_I_K3W(SHHD_EXTERN, double, nt, v, len)
#define SHHD_D_COLWS_F32(nt, v, len) SHHD_I_K3W(SHHD_EXTERN, float, nt, v, len)
#define SHHD_D_ROWWS_F64(nt, v) SHHD_FOR_MAXN(SHHD_I_K2W, SHHD_EXTERN, double, nt, v)
#define SHHD_D_ROWWS_F32(nt, v) SHHD_FOR_MAXN(SHHD_I_K2W, SHHD_EXTERN, float, nt, v)
#ifndef SHHD_DEFINE_INSTANCES
SHHD_FOR_NT_V(SHHD_D_SHAPE_F64)
SHHD_FOR_NT_V(SHHD_D_SHAPE_F32)
SHHD_FOR_ROW_FL(SHHD_D_ROWFL_F64)
SHHD_FOR_ROW_FL(SHHD_D_ROWFL_F32)
SHHD_FOR_COL_FZ(SHHD_D_COLFZ_F64)
SHHD_FOR_COL_FZ(SHHD_D_COLFZ_F32)
SHHD_FOR_COL_FZ_WS(SHHD_D_COLWS_F64)
SHHD_FOR_COL_FZ_WS(SHHD_D_COLWS_F32)
SHHD_FOR_ROW_WS(SHHD_D_ROWWS_F64)
SHHD_FOR_ROW_WS(SHHD_D_ROWWS_F32)
#endif

}  // namespace shhd
