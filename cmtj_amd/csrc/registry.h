// registry.h -- kernel registry shared by the instantiation units.
#pragma once

#include "kernels_cl.h"

namespace cmtj {

struct KEntry {
  int prec;
  int lm;   // layer capacity
  int lpi;  // lanes per instance: 1 (k_junction), lm (k_junction_ll) or 4 lm (k_junction_cl)
  int solver;
  uint32_t fm;
  bool loge;
  const void* fn;
  const char* name;
  int nd1;  // 1 + the number of time-dependent records the kernel is compiled for; 0: any
  int logf;  // 0: any log layout; 1: m_x of every layer + classic R, binary32 entries
};

struct KTable {
  const KEntry* e;
  size_t n;
};

constexpr uint32_t FM_LEAN = F_ANIS | F_IEC | F_DYNH;
constexpr uint32_t FM_STT = F_ANIS | F_IEC | F_DYNH | F_STT;
constexpr uint32_t FM_SOTT = F_ANIS | F_IEC | F_DYN | F_DYNH | F_SOT | F_THERMAL;
// one component per lane, binary64: every term of the field and torque
// (full demag and dipole tensors, second-order anisotropy, biquadratic IEC,
// IDMI, STT, SOT with a secondary reference) and time-dependent drivers
constexpr uint32_t FM_STTC = F_ANIS | F_ANIS2 | F_IEC | F_IEC2 | F_DEMAG_OFF | F_IDMI | F_DIPOLE | F_DYNH | F_DYN |
                             F_STT | F_SOT | F_SOT2;
// coupled runs (stacks, groups): the usual member's features
constexpr uint32_t FM_CPL = F_ANIS | F_IEC | F_DYN | F_DYNH | F_STT | F_SOT | F_THERMAL;

// exact feature sets of the benchmark workloads (see workloads.py)
constexpr uint32_t FM_PIMM = F_ANIS | F_IEC | F_DYNH | F_EXACT | F_KSTEP;          // 2-layer MTJ, Hoe step
constexpr uint32_t FM_RH = F_ANIS | F_EXACT | F_KSTEP;                             // R(H) loop, static drive
constexpr uint32_t FM_VSD = F_ANIS | F_DYNH | F_EXACT | F_KSINE;                   // sine Oersted drive
constexpr uint32_t FM_CIMS = F_ANIS | F_DYN | F_SOT | F_THERMAL | F_EXACT | F_KTRAP;  // SOT trapezoid pulse at T
constexpr uint32_t FM_STO = F_ANIS | F_STT | F_EXACT | F_KSTEP;                    // DC spin-torque oscillator
constexpr uint32_t FM_SOTK = F_ANIS | F_DYN | F_SOT | F_EXACT | F_KSTEP;            // SOT switching by a current step

// one table per instantiation unit
KTable kernels_f64();
KTable kernels_f64e();
KTable kernels_f32();
KTable kernels_spec();
KTable kernels_ll_f64();
KTable kernels_ll_f32();
KTable kernels_ll_spec();
KTable kernels_exact();
KTable kernels_cl();
KTable kernels_cl64();
KTable kernels_stack();  // coupled stacks: chosen by cmtj_run_stacks only
KTable kernels_stack_cl();  // coupled stacks on one lane per component (binary64)
KTable kernels_dp();     // Dormand-Prince (FP64)
const void* dp_queue_kernel();   // single-layer Dormand-Prince from a device work queue
const void* clear_u32_kernel();  // clears a launch's work counters
KTable kernels_group();   // dipole-coupled groups, one wave each: chosen by cmtj_run_groups only
KTable kernels_group4();  // the same on four waves per group

}  // namespace cmtj

#define CMTJ_KE(R, PREC, LM, S, FM, E)                                     \
  {PREC, LM, 1, S, FM, E, (const void*)&cmtj::k_junction<R, LM, S, FM, E>, \
   "k_junction<" #R "," #LM "," #S "," #FM "," #E ">"}

#define CMTJ_KL(R, PREC, L, S, FM)                                               \
  {PREC, L, L, S, FM, false, (const void*)&cmtj::k_junction_ll<R, L, S, FM, false>, \
   "k_junction_ll<" #R "," #L "," #S "," #FM ">"}

// the same with the register allocation capped for WPE waves per SIMD
#define CMTJ_KL_WPE(R, PREC, L, S, FM, WPE)                                              \
  {PREC, L, L, S, FM, false, (const void*)&cmtj::k_junction_ll<R, L, S, FM, false, WPE>, \
   "k_junction_ll<" #R "," #L "," #S "," #FM ",waves " #WPE ">"}

#define CMTJ_KC(L, FM, ND)                                                                     \
  {CMTJ_FP32, L, 4 * (L), CMTJ_RK4, FM, false, (const void*)&cmtj::k_junction_cl<L, FM, ND>,  \
   "k_junction_cl<" #L "," #FM "," #ND ">", (ND) + 1}

#define CMTJ_KCLOG(L, FM, ND)                                                                      \
  {CMTJ_FP32, L, 4 * (L), CMTJ_RK4, FM, false, (const void*)&cmtj::k_junction_cl<L, FM, ND, false, CMTJ_RK4, 1>, \
   "k_junction_cl<" #L "," #FM "," #ND ",log m_x+R f32>", (ND) + 1, 1}

#define CMTJ_KC64(L, FM, ND)                                                                        \
  {CMTJ_FP64, L, 4 * (L), CMTJ_RK4, FM, false, (const void*)&cmtj::k_junction_cl<L, FM, ND, true>,  \
   "k_junction_cl<double," #L "," #FM "," #ND ">", (ND) + 1}

#define CMTJ_KC64EH(L, FM, ND)                                                                     \
  {CMTJ_FP64, L, 4 * (L), CMTJ_EULER_HEUN, FM, false,                                                \
   (const void*)&cmtj::k_junction_cl<L, FM, ND, true, CMTJ_EULER_HEUN>,                              \
   "k_junction_cl<double," #L ",EulerHeun," #FM "," #ND ">", (ND) + 1}

#define CMTJ_KC64H(L, FM, ND)                                                                      \
  {CMTJ_FP64, L, 4 * (L), CMTJ_HEUN, FM, false, (const void*)&cmtj::k_junction_cl<L, FM, ND, true, CMTJ_HEUN>, \
   "k_junction_cl<double," #L ",Heun," #FM "," #ND ">", (ND) + 1}

#define CMTJ_KL_SOLVERS(R, PREC, L, FM) \
  CMTJ_KL(R, PREC, L, CMTJ_RK4, FM), CMTJ_KL(R, PREC, L, CMTJ_HEUN, FM), CMTJ_KL(R, PREC, L, CMTJ_EULER_HEUN, FM)

#define CMTJ_KE_SOLVERS(R, PREC, LM, FM, E)                                     \
  CMTJ_KE(R, PREC, LM, CMTJ_RK4, FM, E), CMTJ_KE(R, PREC, LM, CMTJ_HEUN, FM, E), \
      CMTJ_KE(R, PREC, LM, CMTJ_EULER_HEUN, FM, E)

#define CMTJ_KE_LMS(R, PREC, FM, E)                                                          \
  CMTJ_KE_SOLVERS(R, PREC, 1, FM, E), CMTJ_KE_SOLVERS(R, PREC, 2, FM, E),                   \
      CMTJ_KE_SOLVERS(R, PREC, 4, FM, E), CMTJ_KE_SOLVERS(R, PREC, 8, FM, E)

#define CMTJ_KE_LMS124(R, PREC, FM, E)                                       \
  CMTJ_KE_SOLVERS(R, PREC, 1, FM, E), CMTJ_KE_SOLVERS(R, PREC, 2, FM, E), \
      CMTJ_KE_SOLVERS(R, PREC, 4, FM, E)
