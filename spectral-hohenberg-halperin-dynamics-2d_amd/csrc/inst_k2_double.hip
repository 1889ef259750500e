// Explicit instances of K2 for double.
#define SHHD_DEFINE_INSTANCES
#include "instances.h"
namespace shhd {
#define SHHD_INST_N(nt, v, mn) \
    template __global__ void k_rows<double, nt, v, mn, 0>(const InstDev<double>*, BlockMap);
#define SHHD_INST(nt, v) SHHD_INST_N(nt, v, 0) SHHD_INST_N(nt, v, 1) SHHD_INST_N(nt, v, 2) SHHD_INST_N(nt, v, 4) SHHD_INST_N(nt, v, 8)
SHHD_FOR_NT_V(SHHD_INST)
#define SHHD_INST_P(nt, v) template __global__ void k_rows_physical<double, nt, v>(const InstDev<double>*, double*, int);
SHHD_FOR_NT_V(SHHD_INST_P)
#define SHHD_INST_FN(nt, v, mn, fl) template __global__ void k_rows<double, nt, v, mn, fl>(const InstDev<double>*, BlockMap);
#define SHHD_INST_F(nt, v, fl) SHHD_INST_FN(nt, v, 0, fl) SHHD_INST_FN(nt, v, 1, fl) SHHD_INST_FN(nt, v, 2, fl) \
    SHHD_INST_FN(nt, v, 4, fl) SHHD_INST_FN(nt, v, 8, fl)
SHHD_FOR_ROW_FL(SHHD_INST_F)
}  // namespace shhd
