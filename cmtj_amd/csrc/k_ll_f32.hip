// Generic FP32-mode kernels with one layer per lane, 2/4/8 lanes per instance
// (the only FP32 kernels for 5-8 layers).
#include "registry.h"
namespace cmtj {
static const KEntry kE[] = {CMTJ_KL_SOLVERS(float, CMTJ_FP32, 2, F_ALL), CMTJ_KL_SOLVERS(float, CMTJ_FP32, 4, F_ALL),
                            CMTJ_KL_SOLVERS(float, CMTJ_FP32, 8, F_ALL)};
KTable kernels_ll_f32() { return {kE, sizeof(kE) / sizeof(kE[0])}; }
}  // namespace cmtj
