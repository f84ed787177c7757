// Generic FP64 kernels with the energy channels (calculateEnergies=True).
#include "registry.h"
namespace cmtj {
static const KEntry kE[] = {CMTJ_KE_LMS(double, CMTJ_FP64, F_ALL, true)};
KTable kernels_f64e() { return {kE, sizeof(kE) / sizeof(kE[0])}; }
}  // namespace cmtj
