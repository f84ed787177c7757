// Coupled-stack kernels: one member junction per lane (every feature, FP64),
// the stack's current exchanged between the lanes each step.  Members have
// at most two layers (Stack<T>::isTwoLayerMemberStack, core/stack.hpp:
// 272-283); the reference's stacks step with RK4, or Euler-Heun when a
// member is stochastic (initializeSimulation, :36-62).
#include "registry.h"
namespace cmtj {
constexpr uint32_t FM_STACK = F_ALL;
constexpr uint32_t FM_STT1 = F_ANIS | F_STT | F_EXACT;          // an STT free layer against a fixed polariser
constexpr uint32_t FM_STT2 = F_ANIS | F_IEC | F_STT | F_EXACT;  // the same with a coupled second layer
static const KEntry kS[] = {
    // exact feature sets of the common STT members (feature tests compiled out)
    {CMTJ_FP64, 1, 1, CMTJ_RK4, FM_STT1, false,
     (const void*)&k_junction<double, 1, CMTJ_RK4, FM_STT1, false, true>, "k_junction<double,1,RK4,anis+STT,stack>"},
    {CMTJ_FP64, 2, 1, CMTJ_RK4, FM_STT2, false,
     (const void*)&k_junction<double, 2, CMTJ_RK4, FM_STT2, false, true>,
     "k_junction<double,2,RK4,anis+IEC+STT,stack>"},
    // the usual member: anisotropy, IEC, spin torques, drives, temperature
    {CMTJ_FP64, 1, 1, CMTJ_RK4, FM_CPL, false,
     (const void*)&k_junction<double, 1, CMTJ_RK4, FM_CPL, false, true>, "k_junction<double,1,RK4,coupled,stack>"},
    {CMTJ_FP64, 2, 1, CMTJ_RK4, FM_CPL, false,
     (const void*)&k_junction<double, 2, CMTJ_RK4, FM_CPL, false, true>, "k_junction<double,2,RK4,coupled,stack>"},
    {CMTJ_FP64, 2, 1, CMTJ_EULER_HEUN, FM_CPL, false,
     (const void*)&k_junction<double, 2, CMTJ_EULER_HEUN, FM_CPL, false, true>,
     "k_junction<double,2,EulerHeun,coupled,stack>"},
    // every feature
    {CMTJ_FP64, 2, 1, CMTJ_RK4, FM_STACK, false,
     (const void*)&k_junction<double, 2, CMTJ_RK4, FM_STACK, false, true>, "k_junction<double,2,RK4,stack>"},
    {CMTJ_FP64, 2, 1, CMTJ_EULER_HEUN, FM_STACK, false,
     (const void*)&k_junction<double, 2, CMTJ_EULER_HEUN, FM_STACK, false, true>,
     "k_junction<double,2,EulerHeun,stack>"},
};
KTable kernels_stack() { return {kS, sizeof(kS) / sizeof(kS[0])}; }

// The same stacks with one lane per vector component (4 lanes per layer):
// the exact STT members, the general coupled set (callback drivers read at
// the RK4 stage times) and the stochastic set for Euler-Heun.
#define CMTJ_KSCL(L, SOLVER, FM, NAME)                                                                   \
  {CMTJ_FP64, L, 4 * (L), SOLVER, FM, false,                                                            \
   (const void*)&k_junction_cl<L, FM, -1, true, SOLVER, 0, true>, "k_junction_cl<double," #L "," NAME ",stack>", 0}
static const KEntry kSC[] = {
    CMTJ_KSCL(1, CMTJ_RK4, FM_STT1, "RK4,anis+STT"),
    CMTJ_KSCL(2, CMTJ_RK4, FM_STT2, "RK4,anis+IEC+STT"),
    CMTJ_KSCL(1, CMTJ_RK4, FM_STTC | F_TABLE, "RK4,FM_STTC|F_TABLE"),
    CMTJ_KSCL(2, CMTJ_RK4, FM_STTC | F_TABLE, "RK4,FM_STTC|F_TABLE"),
    CMTJ_KSCL(1, CMTJ_EULER_HEUN, FM_STTC | F_THERMAL, "EulerHeun,FM_STTC|F_THERMAL"),
    CMTJ_KSCL(2, CMTJ_EULER_HEUN, FM_STTC | F_THERMAL, "EulerHeun,FM_STTC|F_THERMAL"),
};
KTable kernels_stack_cl() { return {kSC, sizeof(kSC) / sizeof(kSC[0])}; }
}  // namespace cmtj
