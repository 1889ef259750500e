// Explicit instances of K1 and K3 for float.
#define SHHD_DEFINE_INSTANCES
#include "instances.h"
namespace shhd {
#define SHHD_INST(nt, v) \
    template __global__ void k_inverse_columns<float, nt, v>(const InstDev<float>*, const StagePtrs<float>*, BlockMap); \
    template __global__ void k_forward_columns<float, nt, v, 4>(const InstDev<float>*, const StagePtrs<float>*, StageParams, BlockMap); \
    template __global__ void k_forward_columns<float, nt, v, v>(const InstDev<float>*, const StagePtrs<float>*, StageParams, BlockMap);
SHHD_FOR_NT_V(SHHD_INST)
}  // namespace shhd
