// kernels_cl.h -- the junction kernel with one vector component per lane.
//
// Layer l of an instance occupies the lane quad 4l..4l+3: lanes 0, 1, 2 of
// the quad hold the x, y and z components of every per-layer vector (m, the
// effective field, the torques), lane 3 holds zeros.  An instance of at most
// L layers spans 4L lanes (L = 1, 2 or 4: 4L divides a 16-lane DPP row).
//
// The vector algebra of a step maps onto DPP quad permutations, which the
// compiler folds into the arithmetic instruction that consumes them:
//   (a x b)_c = a_{c+1} b_{c+2} - a_{c+2} b_{c+1}  -- rotations [1,2,0,3] and
//                                                    [2,0,1,3] of the quad
//   a . b      = butterfly sum of a_c b_c over the quad (swaps [1,0,3,2],
//                then [2,3,0,1]); every lane of the quad gets the same sum
// and the neighbour layers (start-of-step values, core/junction.hpp:
// 1682-1707) arrive by a row shift of 4 lanes.  A step of one wave then
// issues about half the instructions of the one-layer-per-lane kernel, and a
// batch occupies four times as many waves: on a batch too small to fill the
// chip both shorten the sweep, whose time is the step count of the longest
// instance times the latency of one wave step.
//
// Integrator: the anchored binary32 RK4 of the other FP32 kernels (rk4_step,
// core/junction.hpp:1017-1041): k1 and the state in binary64, stages 2-4 as
// binary32 increments f(m0 + d, t_s) - f(m0, t) (rhs_delta in kernels.h,
// same terms in component form, the cross products expanded about the
// anchor, so they round differently).  Features: demagnetisation with a diagonal
// tensor, first-order anisotropy, bilinear IEC with constant coupling,
// time-dependent drivers that feed field slots only; log channels m and the
// classic or strip magnetoresistance.  Sums over the quad associate
// differently from the sequential dot products of the other kernels, so
// results agree with them to rounding, not bitwise.
#pragma once

#include <type_traits>

#include "kernels_ll.h"

namespace cmtj {

constexpr int QP_ROT1 = 0xC9;   // quad_perm [1,2,0,3]: component c reads c + 1 (mod 3)
constexpr int QP_ROT2 = 0xD2;   // quad_perm [2,0,1,3]: component c reads c + 2 (mod 3)
constexpr int QP_SWAP1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int QP_SWAP2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int QP_BC0 = 0x00;    // quad_perm [0,0,0,0]: x to every lane of the quad
constexpr int QP_BC1 = 0x55;    // [1,1,1,1]: y
constexpr int QP_BC2 = 0xAA;    // [2,2,2,2]: z
constexpr int DPP_ROW_SHR4 = 0x114;  // lane i reads lane i - 4 (0 outside the row)
constexpr int DPP_ROW_SHL4 = 0x104;  // lane i reads lane i + 4

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
// v of the lane CTRL names, written only to the lanes of the banks (groups
// of 4 lanes in a 16-lane row) that BANK enables; the other lanes keep old.
template <int CTRL, int BANK>
__device__ __forceinline__ double dpp_keep(double old, double v) {
  const unsigned long long o = (unsigned long long)__double_as_longlong(old);
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)(unsigned)(o & 0xffffffffull), (int)(unsigned)(u & 0xffffffffull),
                                             CTRL, 0xF, BANK, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32), CTRL, 0xF, BANK,
                                             false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// Banks whose quad has a layer below (UP = false) or above (UP = true) in
// the same instance, for instances of L = 1, 2 or 4 layers (one quad per
// layer, instances aligned to 4L lanes): the row shifts by 4 lanes then
// never read another instance's lanes, and the lanes of the other banks
// keep the zero they start with.
template <int L, bool UP>
constexpr int neighbour_banks() {
  int m = 0;
  for (int b = 0; b < 4; ++b) {
    const int ly = b % L;
    if (UP ? (ly + 1 < L) : (ly > 0)) m |= 1 << b;
  }
  return m;
}
// Steps between two NaN tests of the component-lane step loop.
constexpr int NAN_BLOCK = 64;

// The smallest nonzero n over the active lanes of the wave (0 if none):
// wave-uniform, read lane by lane through the ballot.
__device__ __forceinline__ uint32_t wave_min_nonzero(uint32_t n) {
  unsigned long long act = __ballot(n != 0u);
  uint32_t best = 0u;
  while (act) {
    const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)n, (int)__builtin_ctzll(act));
    best = (best == 0u || v < best) ? v : best;
    act &= act - 1ull;
  }
  return best;
}
template <int CTRL>
__device__ __forceinline__ double dpp_x(double v) {
  return dpp_d<CTRL>(v);
}
template <int CTRL>
__device__ __forceinline__ float dpp_x(float v) {
  return dpp_f<CTRL>(v);
}

// Sum over the quad, identical in its four lanes: (v0 + v1) + (v2 + v3).
template <class T>
__device__ __forceinline__ T qsum(T v) {
  v = v + dpp_x<QP_SWAP1>(v);
  return v + dpp_x<QP_SWAP2>(v);
}

// (a x b) at this lane's component, given a's rotated components a1 = a_{c+1}
// and a2 = a_{c+2}.
template <class T>
__device__ __forceinline__ T cross_c(T a1, T a2, T b) {
  return a1 * dpp_x<QP_ROT2>(b) - a2 * dpp_x<QP_ROT1>(b);
}

// (a x b) and a x (a x b) at this lane's component, given a's components
// a0 = a_c, a1 = a_{c+1}, a2 = a_{c+2}.  The first product at components
// c + 1 and c + 2 is formed here from the same operands, in the same order,
// as in the lanes that own those components, so it equals their values and
// need not cross lanes before the second product (for binary64 values a
// lane move is two DPP moves and the wait states after them).
template <class T>
__device__ __forceinline__ void cross2_c(T a0, T a1, T a2, T b, T& p, T& q) {
  const T b1 = dpp_x<QP_ROT1>(b), b2 = dpp_x<QP_ROT2>(b);
  p = fma(a1, b2, -(a2 * b1));          // (a x b)_c
  const T p1 = fma(a2, b, -(a0 * b2));  // (a x b)_{c+1}
  const T p2 = fma(a0, b1, -(a1 * b));  // (a x b)_{c+2}
  q = fma(a1, p2, -(a2 * p1));          // (a x (a x b))_c
}

template <class T>
__device__ __forceinline__ T comp(V3<T> v, int c) {
  return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : T(0)));
}

// Whether record k (meta: kind | slot << 8 | layer mask << 16) feeds
// component c of layer ly's field.
__device__ __forceinline__ bool feeds(int meta, int ly, int c) {
  const int type = meta & 0xff;
  const int slot = (meta >> 8) & 0xff;
  const int lmask = (meta >> 16) & 0xff;
  return type != CMTJ_CONSTANT && slot < CMTJ_SLOT_K && slot % 3 == c && ((lmask >> ly) & 1);
}

// Per-lane constants of one layer component.
struct CLConst {
  double hb, an, an1, an2, nms, alpha, gdt, akn, jb, jt, ref;  // binary64 anchor
  float anf, an1f, an2f, nmsf, alphaf, gdtf, aknf;             // binary32 increments
  // binary64 kernels: the layer's scalar drives (constant parts) and spin
  // torques
  double bK, bJt, bJb, bI, bFL, bDL, bT;    // anisotropy, IEC top/bottom, current, SOT field-/damping-like, T
  double bK2, bJ2t, bJ2b;                   // second-order anisotropy, biquadratic IEC top/bottom
  double nd1, nd2;                          // off-diagonal demag row entries (components c + 1, c + 2) x Ms / mu0
  double nb0, nb1, nb2, nt0, nt1, nt2;      // dipole tensor rows (bottom, top) x Ms / mu0
  double bDt, bDb, bFL2, bDL2, ref2;        // IDMI vectors (this component), secondary SOT drives and reference
  double sig2, gconv;                       // 2 alpha kB / (Ms V gamma), -gamma / (1 + alpha^2)
  double inv_ms, inv_mst;                   // 1 / Ms, 1 / (Ms t)
  double sttp, kap, bet, ng, cdt, Pp, lam, lam2;  // gamma hbar / (e Ms t), kappa, beta, -gamma, dt / (1 + a^2), P, L, L^2
  double dlt, flt;                          // static SOT coefficients
  int rt;                                   // reference: none, fixed, top, bottom
  bool stt, alt, sot, dyn_sot, therm;       // Slonczewski torque; alternative form; SOT; SOT from drivers; thermal
};

// The drives of one stage time as a binary64 lane sees them: its field
// component and its layer's scalar slots.
struct CLDrive {
  double H, K, Jt, Jb, I, FL, DL, T, K2, J2t, J2b, FL2, DL2, Dt, Db;
};

// Which scalar slot of layer ly a time-dependent record feeds: 1 anisotropy,
// 2 IEC top, 3 IEC bottom, 4 current, 5 field-like, 6 damping-like torque,
// 7 temperature, 8 second-order anisotropy, 9/10 biquadratic IEC top/bottom,
// 11/12 secondary field-/damping-like torque, 13/14 component c of the IDMI
// vector towards the top/bottom neighbour; 0 none (field slots, other
// slots, other layers, other components).
__device__ __forceinline__ int scalar_slot(int meta, int ly, int c) {
  const int slot = (meta >> 8) & 0xff;
  if ((meta & 0xff) == CMTJ_CONSTANT || !(((meta >> 16) & 0xff) >> ly & 1)) return 0;
  return slot == CMTJ_SLOT_K ? 1 : slot == CMTJ_SLOT_J_TOP ? 2 : slot == CMTJ_SLOT_J_BOTTOM ? 3
       : slot == CMTJ_SLOT_CURRENT ? 4 : slot == CMTJ_SLOT_FL ? 5 : slot == CMTJ_SLOT_DL ? 6
       : slot == CMTJ_SLOT_TEMPERATURE ? 7 : slot == CMTJ_SLOT_K2 ? 8 : slot == CMTJ_SLOT_J2_TOP ? 9
       : slot == CMTJ_SLOT_J2_BOTTOM ? 10 : slot == CMTJ_SLOT_FL2 ? 11 : slot == CMTJ_SLOT_DL2 ? 12
       : slot == CMTJ_SLOT_IDMI_TOP + c ? 13 : slot == CMTJ_SLOT_IDMI_BOTTOM + c ? 14 : 0;
}

__device__ __forceinline__ void scalar_add(CLDrive& d, int kind, double v) {
  d.K += kind == 1 ? v : 0.0;
  d.Jt += kind == 2 ? v : 0.0;
  d.Jb += kind == 3 ? v : 0.0;
  d.I += kind == 4 ? v : 0.0;
  d.FL += kind == 5 ? v : 0.0;
  d.DL += kind == 6 ? v : 0.0;
  d.T += kind == 7 ? v : 0.0;
  d.K2 += kind == 8 ? v : 0.0;
  d.J2t += kind == 9 ? v : 0.0;
  d.J2b += kind == 10 ? v : 0.0;
  d.FL2 += kind == 11 ? v : 0.0;
  d.DL2 += kind == 12 ? v : 0.0;
  d.Dt += kind == 13 ? v : 0.0;
  d.Db += kind == 14 ? v : 0.0;
}

// ND: number of time-dependent records when fixed at compile time (exact
// kernels), -1 to read it from the group.  F64: every stage in binary64
// (the drop-in's arithmetic) instead of the anchored binary32 increments.
// LOGF: 0 reads the log layout from the group; 1 compiles in m_x of every
// layer, the classic resistance and binary32 entries (the PIMM sweep's log).
// STK: the instances are members of electrically coupled stacks
// (Stack<T>::runSimulationKCL / runSimulationNonKCL with delayed coupling,
// core/stack.hpp:305-410): stack s on lanes [s pad, s pad + members x 4L)
// of one wave, member j on the 4L lanes from 4L j; the pair couplings come
// from the members' magnetisations at the start of the step and the stack
// current replaces the members' current at every stage, as in k_junction's
// stack form (binary64 only).
template <int L, uint32_t FM, int ND, bool F64 = false, int SOLVER = CMTJ_RK4, int LOGF = 0, bool STK = false>
__global__ __launch_bounds__(64) void k_junction_cl(GroupArgs g) {
  static_assert(!STK || (F64 && LOGF == 0 && !(FM & F_KSTEP)),
                "stacks: binary64, the group's log layout, every step through the same code");
  static_assert(L == 1 || L == 2 || L == 4, "4L lanes per instance must divide a 16-lane DPP row");
  static_assert(F64 || !(FM & (F_STT | F_SOT | F_SOT2 | F_DYN | F_THERMAL | F_ANIS2 | F_IEC2 | F_DEMAG_OFF |
                                F_IDMI | F_DIPOLE)),
                "binary64 kernels only");
  static_assert(SOLVER == CMTJ_RK4 || (F64 && (SOLVER == CMTJ_EULER_HEUN || SOLVER == CMTJ_HEUN)),
                "RK4, or Euler-Heun / Heun in binary64");
  // callback-driver samples are laid out for RK4's three stage times
  static_assert(!(FM & F_TABLE) || (F64 && SOLVER == CMTJ_RK4), "tabulated drivers: binary64 RK4 only");
  constexpr int LPI = 4 * L;
  const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t inst;
  int q;
  [[maybe_unused]] const cmtj_stack* sk = nullptr;
  [[maybe_unused]] int s_j = 0, s_n = 0, s_base = 0, s_lane0 = 0;
  if constexpr (STK) {
    // lanes of a stack past its members leave at once; a stack never
    // straddles a wave (pad is a power of two of at most 64)
    const uint32_t pad = g.stack_pad, sid = gl / pad;
    if (sid >= g.stack_n) return;
    sk = g.stack + sid;
    s_n = sk->n_junctions;
    s_j = (int)((gl % pad) / LPI);
    if (s_j >= s_n) return;
    q = (int)(gl % LPI);
    inst = sid * (uint32_t)s_n + (uint32_t)s_j;
    s_base = (int)(threadIdx.x & ~(pad - 1));
    s_lane0 = s_base + s_j * LPI;
  } else {
    inst = gl / LPI;
    if (inst >= g.n) return;
    q = (int)(gl % LPI);
  }
  const int ly = q >> 2, c = q & 3;
  const uint32_t fl = g.flags;
  const int nl = g.nl;
  const bool lact = ly < nl;          // the lane's layer exists
  const bool live = lact && c < 3;    // the lane holds a component
  const int lyr = lact ? ly : 0;
  const uint32_t jid = g.order[inst];
  const cmtj_junction* J = g.js + jid;
  const uint32_t nsteps = J->n_steps;
  const uint32_t W = J->write_every;
  const uint32_t NE = J->n_entries;
  const double dt = J->dt;
  const uint64_t lbase = J->log_offset;
  const bool logging = (g.log != nullptr) && g.log_mask != 0;
  const bool f32log = (LOGF == 1) ? true : g.log_f32;
  const uint32_t esize = f32log ? 4u : 8u;
  char* const lp = logging ? (char*)g.log + lbase * esize : nullptr;
  const uint64_t cs = (uint64_t)NE * esize;
  const double Rp = J->Rp, Rap = J->Rap;
  [[maybe_unused]] const uint64_t seed = J->seed, stream = J->stream;
  [[maybe_unused]] const double sdt = sqrt(dt);
  constexpr int sps = 3;
  const bool hasB = lact && ly > 0, hasT = lact && ly + 1 < nl;
  const uint32_t gshift = (threadIdx.x & 63u) & ~(uint32_t)(LPI - 1);  // the instance's first lane in the wave

  // per-lane constants (loadLayer parameters of this layer, this component)
  CLConst K;
  double m;
  {
    const cmtj_layer& Ly = g.layers[J->layer_offset + lyr];
    LP<double> P;
    DV<double> base;
    V3<double> m3;
    load_layer<double>(Ly, g.c, P, base, m3);
    const double z = 0.0;
    m = live ? comp(m3, c) : z;
    K.hb = live ? comp(base.H, c) : z;
    K.an = live ? comp(P.anis, c) : z;
    K.an1 = live ? comp(P.anis, c == 2 ? 0 : c + 1) : z;  // components c + 1, c + 2 (mod 3)
    K.an2 = live ? comp(P.anis, c == 0 ? 2 : c - 1) : z;
    K.nms = live ? comp(mk<double>(P.N[0], P.N[4], P.N[8]), c) * P.ms_mu0 : z;
    K.alpha = P.alpha;
    K.gdt = P.gconv * dt;
    K.akn = live ? K.an * ((2.0 * base.K) * P.inv_ms) : z;  // Ha_c = akn (anis . m)
    K.jb = live ? base.Jb * P.inv_mst : z;
    K.jt = live ? base.Jt * P.inv_mst : z;
    K.ref = live ? comp(mk<double>(Ly.ref[0], Ly.ref[1], Ly.ref[2]), c) : z;
    K.anf = (float)K.an;
    K.an1f = (float)K.an1;
    K.an2f = (float)K.an2;
    K.nmsf = (float)K.nms;
    K.alphaf = (float)K.alpha;
    K.gdtf = (float)P.gconv * (float)dt;
    K.aknf = (float)K.akn;
    K.bK = base.K;
    K.bJt = base.Jt;
    K.bJb = base.Jb;
    K.bI = base.I;
    K.bFL = base.FL;
    K.bDL = base.DL;
    K.bT = base.T;
    K.bK2 = base.K2;
    K.bJ2t = base.J2t;
    K.bJ2b = base.J2b;
    // demag row c against components c + 1 and c + 2 (the rotations)
    K.nd1 = live ? P.N[3 * c + (c == 2 ? 0 : c + 1)] * P.ms_mu0 : z;
    K.nd2 = live ? P.N[3 * c + (c == 0 ? 2 : c - 1)] * P.ms_mu0 : z;
    const int c1 = (c == 2) ? 0 : c + 1, c2 = (c == 0) ? 2 : c - 1;
    K.nb0 = live ? P.Nb[3 * c + c] * P.ms_mu0 : z;
    K.nb1 = live ? P.Nb[3 * c + c1] * P.ms_mu0 : z;
    K.nb2 = live ? P.Nb[3 * c + c2] * P.ms_mu0 : z;
    K.nt0 = live ? P.Nt[3 * c + c] * P.ms_mu0 : z;
    K.nt1 = live ? P.Nt[3 * c + c1] * P.ms_mu0 : z;
    K.nt2 = live ? P.Nt[3 * c + c2] * P.ms_mu0 : z;
    K.bDt = live ? comp(base.Dt, c) : z;
    K.bDb = live ? comp(base.Db, c) : z;
    K.bFL2 = base.FL2;
    K.bDL2 = base.DL2;
    K.ref2 = live ? comp(mk<double>(Ly.ref2[0], Ly.ref2[1], Ly.ref2[2]), c) : z;
    K.sig2 = P.sig2;
    K.gconv = P.gconv;
    K.inv_ms = P.inv_ms;
    K.inv_mst = P.inv_mst;
    const int bits = g.lbits[lyr];
    K.rt = bits & 3;
    K.stt = live && ((bits >> 2) & 3) == 1;
    K.alt = ((bits >> 5) & 1) != 0;
    K.sttp = P.stt_pref;
    K.kap = P.kappa;
    K.bet = P.beta;
    K.ng = P.neg_gamma;
    K.cdt = P.conv * dt;
    K.Pp = P.P;
    K.lam = P.lam;
    K.lam2 = P.lam2;
    K.sot = live && ((bits >> 2) & 3) == 2;
    K.dyn_sot = ((bits >> 4) & 1) != 0;
    K.dlt = P.dl_t;
    K.flt = P.fl_t;
    K.therm = live && ((bits >> 6) & 1) != 0;
  }
  constexpr int NR = (ND >= 0) ? ND : DREG;  // records kept in registers
  static_assert(NR <= DREG, "at most DREG records in registers");
  DynRec dr[DREG];
  bool sel[DREG];
  [[maybe_unused]] int skind[DREG];  // binary64 kernels: the scalar slot each record feeds
  double win_lo = __builtin_inf(), win_hi = -__builtin_inf();
#pragma unroll
  for (int k = 0; k < DREG; ++k) {
    if (k < NR && (ND >= 0 || k < g.ndyn)) {
      dr[k] = load_rec<FM>(g.dyn + (size_t)k * g.n + inst);
      win_lo = fmin(win_lo, dr[k].p1);
      win_hi = fmax(win_hi, dr[k].p2);
      sel[k] = live && feeds(g.dmeta[k], ly, c);
      skind[k] = lact ? scalar_slot(g.dmeta[k], ly, c) : 0;
    } else {
      dr[k] = DynRec{0.0, 0.0, 0.0, 0.0, 0, 0};
      sel[k] = false;
      skind[k] = 0;
    }
  }
  const float sixth_inv = 1.0f / 6.0f;
  // log layout of this lane (logLayerParams, core/junction.hpp:1580-1636):
  // its m channel, if it writes one, and the resistance channels after the
  // m channels of all layers
  const int lmask = logging ? ((LOGF == 1) ? (CMTJ_LOG_MX | CMTJ_LOG_MR) : g.log_mask) : 0;
  const int mrm = (LOGF == 1) ? CMTJ_MR_CLASSIC : g.mr_mode;
  const int nm = ((lmask & CMTJ_LOG_MX) ? 1 : 0) + ((lmask & CMTJ_LOG_MYZ) ? 2 : 0);
  const int mch = (lmask & CMTJ_LOG_MX) ? c : c - 1;
  const bool wm = live && nm && mch >= 0 && ((lmask & CMTJ_LOG_MYZ) || c == 0);
  const uint64_t om = wm ? (uint64_t)(ly * nm + mch) * cs : 0;
  const uint64_t orr = (uint64_t)(nl * nm) * cs;
  const bool classic = (lmask & CMTJ_LOG_MR) && mrm == CMTJ_MR_CLASSIC;
  const bool strip = (lmask & CMTJ_LOG_MR) && mrm == CMTJ_MR_STRIP;
  // compiled-in layout: the lane's one entry per step (R from the pad lane
  // of layer 0), and the vector a single layer's R is taken against (a
  // layer with no layer above holds top = 0 exactly)
  const bool rlane = (LOGF == 1) && classic && q == 3;
  const bool wlog = wm || rlane;
  const uint64_t olog = rlane ? orr : om;
  const double refR = (nl == 1) ? K.ref : 0.0;
  // STRIP coefficients of the lane's layer, read once (the descriptors may
  // sit in host-mapped memory); 0.5 AHE and AMR_Y + SMR_Y are the reference's
  // own first products, so the entries round as before
  double sRx0 = 0.0, sAMRX = 0.0, sSMRX = 0.0, sRy0 = 0.0, sAHE2 = 0.0, sXY = 0.0;
  if (strip && lact && c == 0) {
    sRx0 = J->Rx0[ly];
    sAMRX = J->AMR_X[ly];
    sSMRX = J->SMR_X[ly];
    sRy0 = J->Ry0[ly];
    sAHE2 = 0.5 * J->AHE[ly];
    sXY = J->AMR_Y[ly] + J->SMR_Y[ly];
  }
  uint32_t k = 0;
  int code = CMTJ_OK, nan_layer = -1;
  uint32_t i = 0, iend = nsteps;
  cmtj_state* const S = g.state ? g.state + jid : nullptr;
  if (S) {  // resume: state of the previous slice
    m = live ? S->m[lyr][c] : 0.0;
    i = S->step;
    k = S->entries;
    code = S->code;
    nan_layer = (code == CMTJ_OK) ? -1 : S->nan_layer;
    iend = (code != CMTJ_OK) ? i : (uint32_t)min((uint64_t)nsteps, (uint64_t)i + g.slice);
  }
  uint64_t next_log = (uint64_t)k * W;
  // lanes whose NaN stops a run: component x of every existing layer
  const unsigned long long nan_lanes = __ballot(live && c == 0);
  // neighbour layers' m (start of step): lanes without a neighbour in the
  // instance hold 0 from here on, and the per-step row shifts (dpp_keep)
  // leave them alone.  Lanes of layers the junction does not have (nl < L)
  // hold m = 0 throughout (no field term moves a zero vector), so the
  // existing layer next to them reads 0 from them.
  constexpr int BANK_B = neighbour_banks<L, false>(), BANK_T = neighbour_banks<L, true>();
  double bot = dpp_d<DPP_ROW_SHR4>(m), top = dpp_d<DPP_ROW_SHL4>(m);
  bot = hasB ? bot : 0.0;
  top = hasT ? top : 0.0;
  // stacks: the member's constants and this step's stack current (stage
  // times t, t + dt/2, t + dt) with the stack log's I_j and C_j_{j+1}
  [[maybe_unused]] StackLane sl;
  [[maybe_unused]] int s_ti = 0, s_bi = 0;
  [[maybe_unused]] double sI[3] = {0.0, 0.0, 0.0};
  [[maybe_unused]] double sIlog = 0.0, sClog = 0.0;
  [[maybe_unused]] unsigned long long s_mask = 0ull;  // the stack's lanes
  if constexpr (STK) {
    sl = stack_lane(sk, s_j, s_n);
    s_ti = sk->top[s_j];
    s_bi = sk->bottom[s_j];
    s_mask = ((g.stack_pad >= 64) ? ~0ull : ((1ull << g.stack_pad) - 1)) << (unsigned)s_base;
  }

  // One step of the loop.  DRV: the step's stage times may meet a
  // time-dependent record; without it the driver code is left out (kernels
  // whose records are all step drivers run the steps outside the records'
  // windows that way: those records are exactly 0 there).  LA: every lane
  // of the wave writes a log entry at every step.  `hitb` (wave-uniform)
  // keeps the step's NaN ballot: when it is not 0, only the lanes whose run
  // stops on it have taken the step (their m as the reference leaves it);
  // the others left their state alone and take the step again.
  unsigned long long hitb = 0ull;
  // CHK: the step tests its new m for a NaN (the ballot below); without it
  // the step folds m into nanacc (a NaN stays a NaN in the sum), which the
  // pass tests once per block of steps.
  double nanacc = 0.0;
  auto step = [&](auto drv_tag, auto la_tag, auto chk_tag) __attribute__((always_inline)) {
    constexpr bool DRV = decltype(drv_tag)::value;
    constexpr bool LA = decltype(la_tag)::value;
    constexpr bool CHK = decltype(chk_tag)::value;
    const double t = (double)i * dt;
    double mn;
    if constexpr (STK) {
      // pair couplings from the members' m at the start of the step
      // (storeMagnetisations, core/stack.hpp:65-76): d = m_top . p summed
      // over the member's top-layer quad (in the reference's order), p its
      // bottom layer's m or the polarisation vector
      // (members have one or two layers: the bottom layer is the top layer's
      // neighbour quad, whose start-of-step m this lane holds as top / bot)
      const double pol_c = (c == 0) ? sl.pol[0] : (c == 1) ? sl.pol[1] : (c == 2) ? sl.pol[2] : 0.0;
      double pc = pol_c;
      if constexpr (L > 1) pc = sl.two_layer ? ((s_bi > s_ti) ? top : bot) : pol_c;
      double dq = qsum(m * pc);  // d in the top layer's quad
      if constexpr (L > 1) {
        // and in the member's other quads, by row shifts of 4 lanes
        const double below = dpp_d<DPP_ROW_SHR4>(dq), above = dpp_d<DPP_ROW_SHL4>(dq);
        dq = (ly == s_ti) ? dq : (ly == s_ti + 1) ? below : above;
      }
      stack_current(sl, s_j, s_n, s_base, LPI, dq, t, dt, g.tables, i, sI, sIlog, sClog);
    }
    if constexpr (F64) {
      // rk4_step (core/junction.hpp:1017-1041) in binary64: every stage
      // evaluates the same terms (calculateHeffDipoleInjection, :590-622;
      // solveLLG, :716-815) at its time and magnetisation.  Spin torques use
      // the layer's reference vector: fixed, or the start-of-step top/bottom
      // neighbour; they are computed on every lane and selected per layer,
      // so no quad operation runs under a divergent mask.
      // drives at ts: this component's field slots and the layer's scalar
      // slots, transcendental waveforms in full binary64 (the binary32 path
      // below takes the faster single-precision ones)
      auto drive0 = [&](double ts, int stage) -> CLDrive {
        CLDrive d{K.hb, K.bK, K.bJt, K.bJb, K.bI, K.bFL, K.bDL, K.bT, K.bK2, K.bJ2t, K.bJ2b, K.bFL2, K.bDL2, K.bDt, K.bDb};
        if (!DRV || !(CMTJ_HAS(F_DYNH) || CMTJ_HAS(F_DYN))) return d;
        if constexpr ((FM & F_KSTEP) != 0) {
          // step records are 0 outside [min t_start, max t_stop] (inclusive)
          if (!(ts >= win_lo && ts <= win_hi)) return d;
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          if (ND >= 0 || r < g.ndyn) {
            const double v = wave_value<double, false, FM>(g.dmeta[r] & 0xff, dr[r], g, ts, i, stage, sps);
            d.H = d.H + (sel[r] ? v : 0.0);
            if constexpr ((FM & F_DYN) != 0) scalar_add(d, skind[r], v);
          }
        }
        for (int r = DREG; ND < 0 && !(FM & F_EXACT) && r < g.ndyn; ++r) {
          const DynRec rr = load_rec<FM>(g.dyn + (size_t)r * g.n + inst);
          const double v = wave_value<double, false, FM>(g.dmeta[r] & 0xff, rr, g, ts, i, stage, sps);
          d.H = d.H + ((live && feeds(g.dmeta[r], ly, c)) ? v : 0.0);
          if constexpr ((FM & F_DYN) != 0) scalar_add(d, lact ? scalar_slot(g.dmeta[r], ly, c) : 0, v);
        }
        return d;
      };
      // a stack member's current is the stack's at every stage
      // (executeJunctionStep, core/stack.hpp:78-91)
      auto drive = [&](double ts, int stage) -> CLDrive {
        CLDrive d = drive0(ts, stage);
        if constexpr (STK) d.I = (stage == 0) ? sI[0] : (stage == 1) ? sI[1] : sI[2];
        return d;
      };
      // dm/dt dt at (ms, drives d) with neighbour components tp (top) and
      // bt (bottom)
      auto f = [&](double ms, const CLDrive& d, double tp, double bt) -> double {
        const double s1 = dpp_d<QP_ROT1>(ms), s2 = dpp_d<QP_ROT2>(ms);
        double Hs = d.H;
        const double u = fma(K.an2, s2, fma(K.an1, s1, K.an * ms));  // anis . m
        if (CMTJ_HAS(F_ANIS)) Hs = fma(K.an * ((2.0 * d.K) * K.inv_ms), u, Hs);
        // calculateSecondOrderAnisotropy (core/junction.hpp:645-652)
        if (CMTJ_HAS(F_ANIS2)) Hs = fma(K.an, ((4.0 * d.K2) * (u * u * u)) * K.inv_ms, Hs);
        if (CMTJ_HAS(F_IEC) || CMTJ_HAS(F_IEC2)) {
          // calculateIEC_ (:654-675): J + 2 J2 (m_neighbour . m)
          double st = d.Jt, sb = d.Jb;
          if (CMTJ_HAS(F_IEC2)) {
            st += 2.0 * d.J2t * qsum(tp * ms);
            sb += 2.0 * d.J2b * qsum(bt * ms);
          }
          Hs = fma(tp, st * K.inv_mst, fma(bt, sb * K.inv_mst, Hs));
        }
        if (CMTJ_HAS(F_IDMI)) {
          // calculateIDMI_ (:677-693): -(D x m_neighbour) / (Ms t) per neighbour
          const double dt1 = dpp_d<QP_ROT1>(d.Dt), dt2 = dpp_d<QP_ROT2>(d.Dt);
          const double db1 = dpp_d<QP_ROT1>(d.Db), db2 = dpp_d<QP_ROT2>(d.Db);
          Hs = Hs - (cross_c(db1, db2, bt) + cross_c(dt1, dt2, tp)) * K.inv_mst;
        }
        if (CMTJ_HAS(F_DEMAG_OFF)) Hs = fma(-K.nd1, s1, fma(-K.nd2, s2, Hs));
        Hs = fma(-K.nms, ms, Hs);
        if (CMTJ_HAS(F_DIPOLE)) {
          // calculate_tensor_interaction of the neighbours' dipole tensors
          // (:579-586), with this layer's Ms
          const double b1 = dpp_d<QP_ROT1>(bt), b2 = dpp_d<QP_ROT2>(bt);
          const double t1 = dpp_d<QP_ROT1>(tp), t2 = dpp_d<QP_ROT2>(tp);
          Hs = Hs - (fma(K.nb0, bt, fma(K.nb1, b1, K.nb2 * b2)) + fma(K.nt0, tp, fma(K.nt1, t1, K.nt2 * t2)));
        }
        // m x H and m x (m x H) as cross products: the initial m of a run
        // need not have unit length
        double Pc, Qc;
        cross2_c(ms, s1, s2, Hs, Pc, Qc);
        double r = fma(Qc, K.alpha, Pc) * K.gdt;
        if constexpr ((FM & (F_STT | F_SOT)) != 0) {
          if (CMTJ_HAS(F_STT) || CMTJ_HAS(F_SOT)) {
            const double refc = (K.rt == CMTJ_REF_FIXED) ? K.ref : (K.rt == CMTJ_REF_TOP) ? tp
                              : (K.rt == CMTJ_REF_BOTTOM) ? bt : 0.0;
            double FLc, DLc;  // m x p, m x (m x p)
            cross2_c(ms, s1, s2, refc, FLc, DLc);
            if (CMTJ_HAS(F_SOT)) {
              // SOT (:770-812): coefficients from the drivers, or static
              // ones scaling the current
              const double Hdl = K.dyn_sot ? d.DL : K.dlt * d.I;
              const double Hfl = K.dyn_sot ? d.FL : K.flt * d.I;
              double br = fma(DLc, Hdl + K.alpha * Hfl, fma(FLc, Hfl - K.alpha * Hdl, fma(Qc, K.alpha, Pc)));
              if (CMTJ_HAS(F_SOT2)) {
                // the secondary reference (dynamic SOT drivers only)
                const double Hdl2 = K.dyn_sot ? d.DL2 : 0.0, Hfl2 = K.dyn_sot ? d.FL2 : 0.0;
                double F2, D2;  // m x p2, m x (m x p2)
                cross2_c(ms, s1, s2, K.ref2, F2, D2);
                br = br + F2 * (Hfl2 - K.alpha * Hdl2) + D2 * (Hdl2 + K.alpha * Hfl2);
              }
              r = K.sot ? br * K.gdt : r;
            }
            if (CMTJ_HAS(F_STT)) {
              // Slonczewski STT (:743-769)
              const double mp = qsum(ms * refc);
              const double den = K.alt ? fma(K.lam, mp, 1.0) : fma(K.lam2 - 1.0, mp, K.lam2 + 1.0);
              const double eta = (K.alt ? K.Pp : K.Pp * K.lam2) / den;
              const double stt = (K.sttp * d.I) * eta;
              const double rt =
                  (fma(Qc, K.alpha, Pc) * K.ng + DLc * (-stt * K.kap) + FLc * (stt * K.bet)) * K.cdt;
              r = K.stt ? rt : r;
            }
          }
        }
        return r;
      };
      // stochasticTorque (core/junction.hpp:824-831) with this lane's
      // component dW of the layer's three normals (shared Philox stream)
      [[maybe_unused]] auto stoch_c = [&](double ms, double dW) -> double {
        const double s1 = dpp_d<QP_ROT1>(ms), s2 = dpp_d<QP_ROT2>(ms);
        double c1, c2;  // m x dW, m x (m x dW)
        cross2_c(ms, s1, s2, dW, c1, c2);
        return fma(c2, K.alpha, c1) * K.gconv;
      };
      if constexpr (SOLVER == CMTJ_HEUN) {
        // heunSolverStep (core/junction.hpp:1756-1815): a predictor m + f dt
        // + g sqrt(dt), whose neighbours feed the corrector's second
        // evaluation at t + dt; the same dW in both
        const CLDrive d0 = drive(t, 0);
        const double k1 = f(m, d0, top, bot);
        double dW = 0.0, g1 = 0.0;
        if (CMTJ_HAS(F_THERMAL)) {
          const V3<double> xi = thermal_normals<double>(seed, stream, i, (uint32_t)lyr);
          dW = K.therm ? comp(xi, c) * sqrt(K.sig2 * d0.T) : 0.0;
          g1 = K.therm ? stoch_c(m, dW) : 0.0;
        }
        const double mN = m + k1 + g1 * sdt;
        double bN = 0.0, tN = 0.0;
        if constexpr (L > 1) {
          bN = dpp_keep<DPP_ROW_SHR4, BANK_B>(bot, mN);
          tN = dpp_keep<DPP_ROW_SHL4, BANK_T>(top, mN);
        }
        const double k2 = f(mN, drive(t + dt, 1), tN, bN);
        mn = m + (k1 + k2) * 0.5;
        if (CMTJ_HAS(F_THERMAL)) {
          const double g2 = K.therm ? stoch_c(mN, dW) : 0.0;
          mn = K.therm ? mn + ((g1 + g2) * 0.5) * sdt : mn;
        }
      } else if constexpr (SOLVER == CMTJ_EULER_HEUN) {
        // eulerHeunSolverStep (core/junction.hpp:1709-1754): drift at t,
        // then the stochastic torque g of stochasticTorque (:824-831) at m
        // and at m + g sqrt(dt), with one dW per layer and step (the
        // layer's three normals of the shared Philox stream, this lane's
        // component)
        const CLDrive d0 = drive(t, 0);
        mn = m + f(m, d0, top, bot);
        if (CMTJ_HAS(F_THERMAL)) {
          const V3<double> xi = thermal_normals<double>(seed, stream, i, (uint32_t)lyr);
          const double dW = comp(xi, c) * sqrt(K.sig2 * d0.T);
          const double gn = stoch_c(m, dW);
          const double gp = stoch_c(m + gn * sdt, dW);
          mn = K.therm ? mn + ((gn + gp) * 0.5) * sdt : mn;
        }
      } else {
        const double k1 = f(m, drive(t, 0), top, bot);
        const CLDrive dmid = drive(t + 0.5 * dt, 1);
        const double k2 = f(m + k1 * 0.5, dmid, top, bot);
        const double k3 = f(m + k2 * 0.5, dmid, top, bot);
        const double k4 = f(m + k3, drive(t + dt, 2), top, bot);
        mn = m + (k1 + (k2 * 2.0) + (k3 * 2.0) + k4) / 6.0;
      }
      // CVector::normalize (core/cvector.hpp:185-192): the quad sum adds
      // (x^2 + y^2) + z^2 in the reference's order; a zero vector stays
      const double ss = qsum(mn * mn);
      mn = (ss != 0.0) ? mn / sqrt(ss) : mn;
    } else {
      // ---- anchor: dm/dt at (m, t) in binary64 (calculateHeffDipoleInjection,
      // core/junction.hpp:590-622; solveLLG, :716-815)
      double v0[DREG] = {0.0, 0.0};
      double H = K.hb;
      bool inwin = true;
      if constexpr ((FM & F_KSTEP) != 0) inwin = t >= win_lo && t <= win_hi;
      if (DRV && CMTJ_HAS(F_DYNH) && inwin) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          if (ND >= 0 || r < g.ndyn) {
            v0[r] = wave_value<double, true, FM>(g.dmeta[r] & 0xff, dr[r], g, t, i, 0, sps);
            H = H + (sel[r] ? v0[r] : 0.0);
          }
        }
        for (int r = DREG; ND < 0 && !(FM & F_EXACT) && r < g.ndyn; ++r) {
          const DynRec rr = load_rec<FM>(g.dyn + (size_t)r * g.n + inst);
          const double v = wave_value<double, true, FM>(g.dmeta[r] & 0xff, rr, g, t, i, 0, sps);
          H = H + ((live && feeds(g.dmeta[r], ly, c)) ? v : 0.0);
        }
      }
      // m of the other two components (dot products below sum the three
      // products of each lane's own view: the quad's lanes may differ in the
      // last bit of a dot product, never more)
      const double m1 = dpp_d<QP_ROT1>(m), m2 = dpp_d<QP_ROT2>(m);
      if (CMTJ_HAS(F_ANIS)) H = fma(K.akn, fma(K.an2, m2, fma(K.an1, m1, K.an * m)), H);
      if (CMTJ_HAS(F_IEC)) H = fma(top, K.jt, fma(bot, K.jb, H));
      H = fma(-K.nms, m, H);
      // m x H and m x (m x H) as cross products: the initial m of a run need
      // not have unit length (the reference normalises its constructor
      // argument, not the stored vector), so m (m . H) - H is no substitute
      double P0, Q0;
      cross2_c(m, m1, m2, H, P0, Q0);
      const double k1 = fma(Q0, K.alpha, P0) * K.gdt;

      // ---- stages 2-4: binary32 increments about the anchor
      const float m0a = (float)m1, m0b = (float)m2;
      // the anchor's H and m x H at components c + 1 and c + 2
      const float H0 = (float)H, Pf = (float)P0;
      const float H0a = dpp_f<QP_ROT1>(H0), H0b = dpp_f<QP_ROT2>(H0);
      const float Pfa = dpp_f<QP_ROT1>(Pf), Pfb = dpp_f<QP_ROT2>(Pf);
      auto delta = [&](float d, float ddH) -> float {
        // f(m0 + d, ts) - f(m0, t), component c (rhs_delta, kernels.h),
        // with the cross products expanded about the anchor:
        //   (m0 + d) x (H0 + dH) - m0 x H0 = (m0 + d) x dH + d x H0
        // and the same for m x (m x H) with P0 = m0 x H0, so only the fresh
        // dH and dP cross lanes, and d x H0, d x P0 fill the DPP waits
        const float d1 = dpp_f<QP_ROT1>(d), d2 = dpp_f<QP_ROT2>(d);
        const float s1 = m0a + d1, s2 = m0b + d2;
        float dH = fmaf(-K.nmsf, d, ddH);
        if (CMTJ_HAS(F_ANIS)) dH = fmaf(K.aknf, fmaf(K.an2f, d2, fmaf(K.an1f, d1, K.anf * d)), dH);
        const float dP = cross_c(s1, s2, dH) + fmaf(d1, H0b, -d2 * H0a);
        const float dQ = cross_c(s1, s2, dP) + fmaf(d1, Pfb, -d2 * Pfa);
        return fmaf(dQ, K.alphaf, dP) * K.gdtf;
      };
      auto drive_delta = [&](double ts, int stage) -> float {
        float dd = 0.f;
        if (!DRV || !CMTJ_HAS(F_DYNH)) return dd;
        if constexpr ((FM & F_KSTEP) != 0) {
          if (!(ts >= win_lo && t <= win_hi)) return dd;
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          if (ND >= 0 || r < g.ndyn) {
            const float v = (float)(wave_value<double, true, FM>(g.dmeta[r] & 0xff, dr[r], g, ts, i, stage, sps) - v0[r]);
            dd = dd + (sel[r] ? v : 0.f);
          }
        }
        for (int r = DREG; ND < 0 && !(FM & F_EXACT) && r < g.ndyn; ++r) {
          const DynRec rr = load_rec<FM>(g.dyn + (size_t)r * g.n + inst);
          const int ty = g.dmeta[r] & 0xff;
          const float v = (float)(wave_value<double, true, FM>(ty, rr, g, ts, i, stage, sps) -
                                  wave_value<double, true, FM>(ty, rr, g, t, i, 0, sps));
          dd = dd + ((live && feeds(g.dmeta[r], ly, c)) ? v : 0.f);
        }
        return dd;
      };
      const float dmid = drive_delta(t + 0.5 * dt, 1);
      const float d2 = (float)(k1 * 0.5);
      float dk = delta(d2, dmid);
      float acc = dk * 2.f;
      const float d3 = d2 + dk * 0.5f;
      dk = delta(d3, dmid);
      acc = acc + dk * 2.f;
      const float dend = drive_delta(t + dt, 2);
      const float d4 = (float)k1 + dk;
      acc = acc + delta(d4, dend);
      mn = m + k1 + (double)(acc * sixth_inv);
      // CVector::normalize (core/cvector.hpp:185-192) over the quad: binary32
      // reciprocal square root and one Newton step in binary64 (|m| = 1 to
      // ~2e-14, far inside the increments' binary32 rounding).  The 1e-30
      // added to the binary32 s leaves any s near 1 unchanged and a NaN a
      // NaN, and keeps y finite for a zero vector, which so stays as it is
      const double s = qsum(mn * mn);
      double y = (double)__frsqrt_rn((float)s + 1e-30f);
      y = y * (1.5 - (0.5 * s) * (y * y));
      mn = mn * y;
    }
    // the end of a step that goes on: new m, neighbours, log entry
    auto finish = [&]() __attribute__((always_inline)) {
      m = mn;
      if constexpr (L > 1) {
        bot = dpp_keep<DPP_ROW_SHR4, BANK_B>(bot, m);
        top = dpp_keep<DPP_ROW_SHL4, BANK_T>(top, m);
      }
      if (LA || (uint64_t)i == next_log) {
        char* const p = lp + (uint64_t)k * esize;
        if constexpr (LOGF == 1) {
          // the compiled-in layout (m_x of every layer, then the classic R;
          // binary32 entries): one store per writing lane, m_x from the x
          // lane of each layer, R from the pad lane of layer 0
          // (calculateMagnetoresistance, :1865-1867, and against the fixed
          // reference vector for one layer, :1623-1629)
          const double cosT = qsum(m * (top + refR));
          const double rv = Rp + (((Rap - Rp) / 2.0) * (1.0 - cosT));
          if (wlog) log_put(true, p + olog, rlane ? rv : m);
        } else {
          if (wm) log_put(f32log, p + om, m);
          if (classic) {
            // calculateMagnetoresistance (:1865-1867): layer 0 against layer 1,
            // one layer against the fixed reference vector (:1623-1629)
            const double cosT = qsum(m * (nl == 1 ? K.ref : top));
            if (q == 0) log_put(f32log, p + orr, Rp + (((Rap - Rp) / 2.0) * (1.0 - cosT)));
          } else if (strip) {
            // stripMagnetoResistance (:1835-1857), parallel sum of the layers
            const double mx = dpp_d<QP_BC0>(m), my = dpp_d<QP_BC1>(m), mz = dpp_d<QP_BC2>(m);
            double rx = 0.0, ry = 0.0;
            if (lact && c == 0) {
              const double Rx = sRx0 + sAMRX * (mx * mx) + sSMRX * (my * my);
              const double Ry = sRy0 + sAHE2 * mz + sXY * mx * my;
              rx = 1. / Rx;
              ry = 1. / Ry;
            }
#pragma unroll
            for (int o = 4; o < LPI; o <<= 1) {
              rx += __shfl_xor(rx, o, 64);
              ry += __shfl_xor(ry, o, 64);
            }
            if (q == 0) {
              log_put(f32log, p + orr, 1 / rx);
              log_put(f32log, p + orr + cs, 1 / ry);
            }
          }
        }
        if constexpr (STK) {
          if (logging && q == 0) {  // the member's stack channels follow its junction channels
            const uint64_t nch = (uint64_t)log_channels(nl, g.mr_mode, g.log_mask);
            log_put(f32log, p + nch * cs, sIlog);
            log_put(f32log, p + (nch + 1) * cs, sClog);
          }
        }
        ++k;
        if constexpr (!LA) next_log += W;
      }
    };
    // the reference updates layers in turn and stops at the first NaN
    // layer; layers up to and including it take the new value (RK4 only:
    // the Euler-Heun step of the reference has no such test)
    if constexpr (!CHK) {
      if constexpr (SOLVER == CMTJ_RK4) nanacc = nanacc + mn;
      finish();
      return;
    }
    const unsigned long long bad = (SOLVER == CMTJ_RK4) ? (__ballot(isnan(mn)) & nan_lanes) : 0ull;
    hitb = bad;
    if (__builtin_expect(bad != 0ull, 0)) {
      // a NaN in the wave (rare): layers up to and including the instance's
      // first NaN layer take the new value; a member's NaN ends its stack's
      // run (the reference throws), the other members take this step and
      // stop with it (selects, no per-lane branch: the whole test stays
      // wave-uniform)
      const uint32_t gb = (uint32_t)((bad >> gshift) & ((1ull << LPI) - 1));
      bool take = (gb != 0u) & (ly <= (int)(__builtin_ctz(gb | 0x80000000u) >> 2));
      if constexpr (STK) take = take | ((gb == 0u) & ((bad & s_mask) != 0ull));
      m = take ? mn : m;
      return;
    }
    finish();  // the common case, wave-uniform: no NaN in the wave
  };
  bool stopped = false;
  // Every lane logs every step when each has W = 1 and its next entry due
  // now (then entries and steps advance together).
  const bool log_all = __ballot(!(W == 1u && next_log == (uint64_t)i)) == 0ull;
  // Steps of one phase (the lane's steps below lim) in wave-uniform passes:
  // each pass runs the fewest steps any lane of the wave still has in the
  // phase on a scalar counter, so the loop needs no per-lane exit test;
  // lanes done with the phase sit out the passes that follow.  A NaN in any
  // lane ends the pass at that step, which no lane counts: the lanes of
  // the instances (stacks) it stops end their run there (steps done = i),
  // the others take the step again in the next pass (the same inputs: the
  // same result).
  auto phase = [&](uint32_t lim, auto drv_tag, auto la_tag) __attribute__((always_inline)) {
    for (;;) {
      const uint32_t n = (!stopped && i < lim) ? lim - i : 0u;
      const uint32_t seg = wave_min_nonzero(n);
      if (seg == 0u) break;
      if (n != 0u) {
        const uint32_t i0 = i;
        uint32_t end = seg;  // wave-uniform
        if constexpr (SOLVER == CMTJ_RK4) {
          // blocks of NAN_BLOCK steps without a per-step NaN test; a block
          // whose sum shows a NaN is taken again from its start, step by
          // step with the test, up to the NaN step (the same inputs: the
          // same values; log entries a stopping instance wrote past that
          // step lie beyond its entries_done)
          for (uint32_t u0 = 0; u0 < end;) {
            const uint32_t ue = min(u0 + (uint32_t)NAN_BLOCK, end);
            const double m_b = m, bot_b = bot, top_b = top;
            const uint32_t k_b = k;
            const uint64_t nl_b = next_log;
            nanacc = 0.0;
            for (uint32_t u = u0; u < ue; ++u) {
              i = i0 + u;
              step(drv_tag, la_tag, std::false_type{});
            }
            if (__builtin_expect((__ballot(isnan(nanacc)) & nan_lanes) != 0ull, 0)) {
              m = m_b;
              bot = bot_b;
              top = top_b;
              k = k_b;
              next_log = nl_b;
              for (uint32_t u = u0; u < ue; ++u) {
                i = i0 + u;
                step(drv_tag, la_tag, std::true_type{});
                if (hitb) {
                  end = u;
                  break;
                }
              }
            }
            u0 = ue;
          }
        } else {
          for (uint32_t u = 0; u < end; ++u) {
            i = i0 + u;
            step(drv_tag, la_tag, std::false_type{});
          }
        }
        i = i0 + end;
        if (hitb) {
          const uint32_t gb = (uint32_t)((hitb >> gshift) & ((1ull << LPI) - 1));
          bool stop = gb != 0u;
          if constexpr (STK) stop = stop || (hitb & s_mask) != 0ull;
          if (gb) {
            nan_layer = __builtin_ctz(gb) >> 2;
            code = CMTJ_ERR_NAN;
          }
          // every lane whose NaN is in the ballot belongs to an instance
          // that stops here, so the pass makes progress; were none to stop,
          // the lanes would take the same step again for ever: end their
          // runs with the error instead
          if (__ballot(stop) == 0ull) {
            code = CMTJ_ERR_NAN;
            stop = true;
          }
          stopped = stop;
          hitb = 0ull;
        }
      }
    }
  };
  auto run = [&](uint32_t lim, auto drv_tag) __attribute__((always_inline)) {
    if (log_all)
      phase(lim, drv_tag, std::true_type{});
    else
      phase(lim, drv_tag, std::false_type{});
  };
  if constexpr ((FM & F_KSTEP) != 0 && !(FM & F_TABLE)) {
    // steps [ia, ib) may meet a record's window [win_lo, win_hi]; every
    // stage time of a step before ia is below win_lo, of a step from ib on
    // above win_hi (one step of margin on each side)
    uint32_t ia = iend, ib = iend;
    if (win_lo <= win_hi) {
      const double a = floor(win_lo / dt) - 2.0, b = floor(win_hi / dt) + 2.0;
      ia = a <= (double)i ? i : (a >= (double)iend ? iend : (uint32_t)a);
      ib = b <= (double)ia ? ia : (b >= (double)iend ? iend : (uint32_t)b);
    }
    run(ia, std::false_type{});
    run(ib, std::true_type{});
    run(iend, std::false_type{});
  } else {
    run(iend, std::true_type{});
  }
  if (g.mag_out && live) g.mag_out[((size_t)jid * CMTJ_MAX_LAYERS + ly) * 3 + c] = m;
  if (g.status && q == 0) {
    cmtj_status s;
    s.code = code;
    s.steps_done = i;
    s.entries_done = k;
    s.nan_layer = nan_layer;
    g.status[jid] = s;
  }
  if (S) {
    if (live) S->m[ly][c] = m;
    if (q == 0) {
      S->step = i;
      S->entries = k;
      S->code = code;
      S->nan_layer = nan_layer;
    }
  }
}

}  // namespace cmtj
