// Explicit instances of K1 and K3 for double.
#define SHHD_DEFINE_INSTANCES
#include "instances.h"
namespace shhd {
#define SHHD_INST(nt, v) \
    template __global__ void k_inverse_columns<double, nt, v>(const InstDev<double>*, const StagePtrs<double>*, BlockMap); \
    template __global__ void k_forward_columns<double, nt, v, 4>(const InstDev<double>*, const StagePtrs<double>*, StageParams, BlockMap); \
    template __global__ void k_forward_columns<double, nt, v, v>(const InstDev<double>*, const StagePtrs<double>*, StageParams, BlockMap);
SHHD_FOR_NT_V(SHHD_INST)
}  // namespace shhd
