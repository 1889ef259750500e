// Explicit instances of K1 and K3 for double.
#define SHHD_DEFINE_INSTANCES
#include "instances.h"
namespace shhd {
#define SHHD_INST(nt, v) \
    template __global__ void k_inverse_columns<double, nt, v>(Desc, StagePtrs<double>, cx<double>*, const cx<double>*, int); \
    template __global__ void k_forward_columns<double, nt, v, 4>(Desc, StagePtrs<double>, StageParams, const cx<double>*, const cx<double>*, int); \
    template __global__ void k_forward_columns<double, nt, v, v>(Desc, StagePtrs<double>, StageParams, const cx<double>*, const cx<double>*, int);
SHHD_FOR_NT_V(SHHD_INST)
}  // namespace shhd
