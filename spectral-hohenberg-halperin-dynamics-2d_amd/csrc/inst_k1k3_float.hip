// Explicit instances of the staged K1 and K3 and of K2 (any length, and the
// physical-field row pass) for float.
#define SHHD_DEFINE_INSTANCES
#include "instances.h"
namespace shhd {
#define SHHD_B(nt, v) SHHD_I_SHAPE(template, float, nt, v)
SHHD_FOR_NT_V(SHHD_B)
}  // namespace shhd
