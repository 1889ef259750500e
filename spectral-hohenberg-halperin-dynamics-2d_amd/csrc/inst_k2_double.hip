// Explicit instances of the fixed-length K2 builds for double.
#define SHHD_DEFINE_INSTANCES
#include "instances.h"
namespace shhd {
#define SHHD_B(nt, v, fl) SHHD_I_ROWFL(template, double, nt, v, fl)
SHHD_FOR_ROW_FL(SHHD_B)
}  // namespace shhd
