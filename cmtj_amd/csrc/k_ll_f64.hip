// Generic FP64 kernels with one layer per lane, 2/4/8 lanes per instance.
#include "registry.h"
namespace cmtj {
static const KEntry kE[] = {CMTJ_KL_SOLVERS(double, CMTJ_FP64, 2, F_ALL), CMTJ_KL_SOLVERS(double, CMTJ_FP64, 4, F_ALL),
                            CMTJ_KL_SOLVERS(double, CMTJ_FP64, 8, F_ALL)};
KTable kernels_ll_f64() { return {kE, sizeof(kE) / sizeof(kE[0])}; }
}  // namespace cmtj
