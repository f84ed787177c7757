// Kernels specialised to common feature sets: fewer live registers and no
// code for absent terms.
#include "registry.h"
namespace cmtj {
static const KEntry kE[] = {
    // anisotropy + IEC + time-dependent drivers (R(H), PIMM, FMR sweeps)
    CMTJ_KE(double, CMTJ_FP64, 1, CMTJ_RK4, FM_LEAN, false),
    CMTJ_KE(double, CMTJ_FP64, 2, CMTJ_RK4, FM_LEAN, false),
    CMTJ_KE(float, CMTJ_FP32, 1, CMTJ_RK4, FM_LEAN, false),
    CMTJ_KE(float, CMTJ_FP32, 2, CMTJ_RK4, FM_LEAN, false),
    // spin-transfer oscillators
    CMTJ_KE(double, CMTJ_FP64, 1, CMTJ_RK4, FM_STT, false),
    CMTJ_KE(double, CMTJ_FP64, 2, CMTJ_RK4, FM_STT, false),
    CMTJ_KE(float, CMTJ_FP32, 1, CMTJ_RK4, FM_STT, false),
    CMTJ_KE(float, CMTJ_FP32, 2, CMTJ_RK4, FM_STT, false),
    // stochastic spin-orbit switching
    CMTJ_KE(double, CMTJ_FP64, 1, CMTJ_EULER_HEUN, FM_SOTT, false),
    CMTJ_KE(double, CMTJ_FP64, 2, CMTJ_EULER_HEUN, FM_SOTT, false),
    CMTJ_KE(float, CMTJ_FP32, 1, CMTJ_EULER_HEUN, FM_SOTT, false),
    CMTJ_KE(float, CMTJ_FP32, 2, CMTJ_EULER_HEUN, FM_SOTT, false),
    CMTJ_KE(float, CMTJ_FP32, 1, CMTJ_HEUN, FM_SOTT, false),
    CMTJ_KE(float, CMTJ_FP32, 2, CMTJ_HEUN, FM_SOTT, false),
};
KTable kernels_spec() { return {kE, sizeof(kE) / sizeof(kE[0])}; }
}  // namespace cmtj
