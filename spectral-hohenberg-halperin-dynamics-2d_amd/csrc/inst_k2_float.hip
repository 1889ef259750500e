// Explicit instances of the fixed-length K2 builds for float.
#define SHHD_DEFINE_INSTANCES
#include "instances.h"
namespace shhd {
#define SHHD_B(nt, v, fl) SHHD_I_ROWFL(template, float, nt, v, fl)
SHHD_FOR_ROW_FL(SHHD_B)
#define SHHD_W(nt, v) SHHD_FOR_MAXN(SHHD_I_K2W, template, float, nt, v)
SHHD_FOR_ROW_WS(SHHD_W)
}  // namespace shhd
