// Explicit instances of K1 and K3 for float.
#define SHHD_DEFINE_INSTANCES
#include "instances.h"
namespace shhd {
#define SHHD_INST(nt, v) \
    template __global__ void k_inverse_columns<float, nt, v>(Desc, StagePtrs<float>, cx<float>*, const cx<float>*, int); \
    template __global__ void k_forward_columns<float, nt, v, 4>(Desc, StagePtrs<float>, StageParams, const cx<float>*, const cx<float>*, int); \
    template __global__ void k_forward_columns<float, nt, v, v>(Desc, StagePtrs<float>, StageParams, const cx<float>*, const cx<float>*, int);
SHHD_FOR_NT_V(SHHD_INST)
}  // namespace shhd
