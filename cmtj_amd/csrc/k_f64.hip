// Generic FP64 kernels (every feature compiled in), 1/2/4/8 layers.
#include "registry.h"
namespace cmtj {
static const KEntry kE[] = {CMTJ_KE_LMS(double, CMTJ_FP64, F_ALL, false)};
KTable kernels_f64() { return {kE, sizeof(kE) / sizeof(kE[0])}; }
}  // namespace cmtj
