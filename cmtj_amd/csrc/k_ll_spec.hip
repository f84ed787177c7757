// One-layer-per-lane kernels specialised to the feature sets of the common
// two-layer workloads.
#include "registry.h"
namespace cmtj {
static const KEntry kE[] = {
    CMTJ_KL(double, CMTJ_FP64, 2, CMTJ_RK4, FM_LEAN),
    CMTJ_KL(float, CMTJ_FP32, 2, CMTJ_RK4, FM_LEAN),
    CMTJ_KL(double, CMTJ_FP64, 2, CMTJ_RK4, FM_STT),
    CMTJ_KL(float, CMTJ_FP32, 2, CMTJ_RK4, FM_STT),
    CMTJ_KL(double, CMTJ_FP64, 2, CMTJ_EULER_HEUN, FM_SOTT),
    CMTJ_KL(float, CMTJ_FP32, 2, CMTJ_EULER_HEUN, FM_SOTT),
    CMTJ_KL(float, CMTJ_FP32, 2, CMTJ_HEUN, FM_SOTT),
};
KTable kernels_ll_spec() { return {kE, sizeof(kE) / sizeof(kE[0])}; }
}  // namespace cmtj
