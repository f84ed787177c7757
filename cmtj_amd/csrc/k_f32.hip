// Generic FP32-bulk kernels (every feature compiled in), 1/2/4 layers.
#include "registry.h"
namespace cmtj {
static const KEntry kE[] = {CMTJ_KE_LMS124(float, CMTJ_FP32, F_ALL, false)};
KTable kernels_f32() { return {kE, sizeof(kE) / sizeof(kE[0])}; }
}  // namespace cmtj
