// Kernels for the exact feature sets and driver kinds of the benchmark
// workloads: every feature test and the driver-kind dispatch resolve at
// compile time.
#include "registry.h"
namespace cmtj {
static const KEntry kE[] = {
    CMTJ_KL(float, CMTJ_FP32, 2, CMTJ_RK4, FM_PIMM),
    CMTJ_KL(double, CMTJ_FP64, 2, CMTJ_RK4, FM_PIMM),
    CMTJ_KL(float, CMTJ_FP32, 4, CMTJ_RK4, FM_PIMM),   // 3-4 layer IEC chains with a field step
    CMTJ_KL(double, CMTJ_FP64, 4, CMTJ_RK4, FM_PIMM),
    CMTJ_KE(float, CMTJ_FP32, 2, CMTJ_RK4, FM_PIMM, false),
    CMTJ_KE(double, CMTJ_FP64, 2, CMTJ_RK4, FM_PIMM, false),
    CMTJ_KE(float, CMTJ_FP32, 1, CMTJ_RK4, FM_RH, false),
    CMTJ_KE(double, CMTJ_FP64, 1, CMTJ_RK4, FM_RH, false),
    CMTJ_KE(float, CMTJ_FP32, 1, CMTJ_RK4, FM_VSD, false),
    CMTJ_KE(double, CMTJ_FP64, 1, CMTJ_RK4, FM_VSD, false),
    CMTJ_KL_WPE(float, CMTJ_FP32, 2, CMTJ_EULER_HEUN, FM_CIMS, 4),  // the ensemble fills the chip
    CMTJ_KL(double, CMTJ_FP64, 2, CMTJ_EULER_HEUN, FM_CIMS),
    CMTJ_KE(float, CMTJ_FP32, 2, CMTJ_EULER_HEUN, FM_CIMS, false),
    CMTJ_KL(float, CMTJ_FP32, 2, CMTJ_RK4, FM_STO),
    CMTJ_KL(double, CMTJ_FP64, 2, CMTJ_RK4, FM_STO),
};
KTable kernels_exact() { return {kE, sizeof(kE) / sizeof(kE[0])}; }
}  // namespace cmtj
