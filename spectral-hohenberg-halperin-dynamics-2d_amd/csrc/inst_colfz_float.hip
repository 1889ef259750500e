// Explicit instances of the fused K1 and K3 builds for float.
#define SHHD_DEFINE_INSTANCES
#include "instances.h"
namespace shhd {
#define SHHD_B(nt, v, len) SHHD_I_COLFZ(template, float, nt, v, len)
SHHD_FOR_COL_FZ(SHHD_B)
#define SHHD_W(nt, v, len) SHHD_I_K3W(template, float, nt, v, len)
SHHD_FOR_COL_FZ_WS(SHHD_W)
}  // namespace shhd
