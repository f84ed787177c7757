// Dipole-coupled group kernels: one group per workgroup, one element per
// lane of its first wave (every feature, FP64), the group field of each step
// built in LDS from the elements' top-layer magnetisations (GroupInteraction,
// core/reservoir.hpp:121-230).  The reference steps a group with RK4, or
// with Euler-Heun for elements that are stochastic (getSolver,
// core/junction.hpp:1892-1931, called from runSimulation, reservoir.hpp:
// 192-212).  Two tables: one wave per group (many groups fill the chip) and
// four waves per group, the sum over the sources split four ways (a few
// groups: shorter steps).
#include "registry.h"
namespace cmtj {
constexpr uint32_t FM_GROUP = F_ALL;

#define CMTJ_KG(L, S, FM, GW, NAME) \
  {CMTJ_FP64, L, 1, S, FM, false, (const void*)&k_group<double, L, S, FM, GW>, NAME}

static const KEntry kG[] = {
    // the usual reservoir element: anisotropy, IEC, field drives
    CMTJ_KG(1, CMTJ_RK4, FM_LEAN, 1, "k_group<double,1,RK4,lean,1>"),
    CMTJ_KG(2, CMTJ_RK4, FM_LEAN, 1, "k_group<double,2,RK4,lean,1>"),
    // spin torques, scalar drives, temperature
    CMTJ_KG(2, CMTJ_RK4, FM_CPL, 1, "k_group<double,2,RK4,coupled,1>"),
    CMTJ_KG(2, CMTJ_EULER_HEUN, FM_CPL, 1, "k_group<double,2,EulerHeun,coupled,1>"),
    // every feature
    CMTJ_KG(2, CMTJ_RK4, FM_GROUP, 1, "k_group<double,2,RK4,1>"),
    CMTJ_KG(4, CMTJ_RK4, FM_GROUP, 1, "k_group<double,4,RK4,1>"),
    CMTJ_KG(2, CMTJ_EULER_HEUN, FM_GROUP, 1, "k_group<double,2,EulerHeun,1>"),
};
static const KEntry kG4[] = {
    CMTJ_KG(1, CMTJ_RK4, FM_LEAN, 4, "k_group<double,1,RK4,lean,4>"),
    CMTJ_KG(2, CMTJ_RK4, FM_LEAN, 4, "k_group<double,2,RK4,lean,4>"),
    CMTJ_KG(2, CMTJ_RK4, FM_CPL, 4, "k_group<double,2,RK4,coupled,4>"),
    CMTJ_KG(2, CMTJ_EULER_HEUN, FM_CPL, 4, "k_group<double,2,EulerHeun,coupled,4>"),
    CMTJ_KG(2, CMTJ_RK4, FM_GROUP, 4, "k_group<double,2,RK4,4>"),
    CMTJ_KG(4, CMTJ_RK4, FM_GROUP, 4, "k_group<double,4,RK4,4>"),
    CMTJ_KG(2, CMTJ_EULER_HEUN, FM_GROUP, 4, "k_group<double,2,EulerHeun,4>"),
};
KTable kernels_group() { return {kG, sizeof(kG) / sizeof(kG[0])}; }
KTable kernels_group4() { return {kG4, sizeof(kG4) / sizeof(kG4[0])}; }
}  // namespace cmtj
