// Dipole-coupled group kernels: one group per workgroup, one element per
// lane (every feature, FP64), the group field of each step built in LDS from
// the elements' top-layer magnetisations (GroupInteraction,
// core/reservoir.hpp:121-230).  The reference steps a group with RK4, or
// with Euler-Heun for elements that are stochastic (getSolver,
// core/junction.hpp:1892-1931, called from runSimulation, reservoir.hpp:
// 192-212).
#include "registry.h"
namespace cmtj {
constexpr uint32_t FM_GROUP = F_ALL;
static const KEntry kG[] = {
    // the usual reservoir element: anisotropy, demagnetisation, fields
    {CMTJ_FP64, 1, 1, CMTJ_RK4, FM_LEAN, false,
     (const void*)&k_junction<double, 1, CMTJ_RK4, FM_LEAN, false, false, true>, "k_junction<double,1,RK4,lean,group>"},
    {CMTJ_FP64, 2, 1, CMTJ_RK4, FM_LEAN, false,
     (const void*)&k_junction<double, 2, CMTJ_RK4, FM_LEAN, false, false, true>, "k_junction<double,2,RK4,lean,group>"},
    // spin torques, scalar drives, temperature
    {CMTJ_FP64, 2, 1, CMTJ_RK4, FM_CPL, false,
     (const void*)&k_junction<double, 2, CMTJ_RK4, FM_CPL, false, false, true>,
     "k_junction<double,2,RK4,coupled,group>"},
    {CMTJ_FP64, 2, 1, CMTJ_EULER_HEUN, FM_CPL, false,
     (const void*)&k_junction<double, 2, CMTJ_EULER_HEUN, FM_CPL, false, false, true>,
     "k_junction<double,2,EulerHeun,coupled,group>"},
    // every feature
    {CMTJ_FP64, 2, 1, CMTJ_RK4, FM_GROUP, false,
     (const void*)&k_junction<double, 2, CMTJ_RK4, FM_GROUP, false, false, true>, "k_junction<double,2,RK4,group>"},
    {CMTJ_FP64, 4, 1, CMTJ_RK4, FM_GROUP, false,
     (const void*)&k_junction<double, 4, CMTJ_RK4, FM_GROUP, false, false, true>, "k_junction<double,4,RK4,group>"},
    {CMTJ_FP64, 2, 1, CMTJ_EULER_HEUN, FM_GROUP, false,
     (const void*)&k_junction<double, 2, CMTJ_EULER_HEUN, FM_GROUP, false, false, true>,
     "k_junction<double,2,EulerHeun,group>"},
};
KTable kernels_group() { return {kG, sizeof(kG) / sizeof(kG[0])}; }
}  // namespace cmtj
