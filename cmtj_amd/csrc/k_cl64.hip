// One-component-per-lane kernels in binary64 (every RK4 stage, the
// drop-in's arithmetic): the feature sets of the binary32 ones, and a set
// with spin torques and time-dependent scalar drives.
#include "registry.h"
namespace cmtj {
static const KEntry kC64[] = {
    CMTJ_KC64(2, FM_PIMM, 1),  CMTJ_KC64(2, FM_PIMM, -1), CMTJ_KC64(4, FM_PIMM, -1), CMTJ_KC64(1, FM_RH, 0),
    CMTJ_KC64(1, FM_VSD, 1),   CMTJ_KC64(1, FM_LEAN, -1), CMTJ_KC64(2, FM_LEAN, -1), CMTJ_KC64(4, FM_LEAN, -1),
    // spin torques, scalar drives: the STO set and the lean set with both
    CMTJ_KC64(2, FM_STO, -1),
    // RK4 also reads callback drivers sampled at its three stage times
    CMTJ_KC64(1, FM_STTC | F_TABLE, -1), CMTJ_KC64(2, FM_STTC | F_TABLE, -1), CMTJ_KC64(4, FM_STTC | F_TABLE, -1),
    // SOT switching by a current step, deterministic and at temperature
    CMTJ_KC64(1, FM_SOTK, 1), CMTJ_KC64(2, FM_SOTK, 1), CMTJ_KC64EH(1, FM_CIMS, 1), CMTJ_KC64EH(2, FM_CIMS, 1),
    // Euler-Heun (stochastic runs): the same set with temperature
    CMTJ_KC64EH(1, FM_STTC | F_THERMAL, -1), CMTJ_KC64EH(2, FM_STTC | F_THERMAL, -1),
    CMTJ_KC64EH(4, FM_STTC | F_THERMAL, -1),
    // Heun
    CMTJ_KC64H(1, FM_STTC | F_THERMAL, -1), CMTJ_KC64H(2, FM_STTC | F_THERMAL, -1),
    CMTJ_KC64H(4, FM_STTC | F_THERMAL, -1),
};
KTable kernels_cl64() { return {kC64, sizeof(kC64) / sizeof(kC64[0])}; }
}  // namespace cmtj
