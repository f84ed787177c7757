// kernels.h -- the junction integration kernel family.
//
// One junction instance per lane.  All layers of an instance live in the
// lane's registers; per step the kernel runs the reference's integrator
// (RK4 with start-of-step neighbours, Heun, Euler-Heun: core/junction.hpp:
// 1017-1041, 1682-1707, 1709-1815) and writes a decimated log entry every
// write_every steps (core/junction.hpp:1973-1981).
//
// Template parameters: R (double: the FP64 drop-in; float: field/torque
// stages in binary32 with state m, clock and driver phase in binary64), LM
// (layer capacity), SOLVER, FM (compiled-in features; the batch-uniform
// runtime flags select among them), LOGE (energy channels).
#pragma once

#include "physics.h"

namespace cmtj {

constexpr int MAX_DYN = 16;

// Numbers of one time-dependent driver of one lane, stored [k][lane] so that
// a wave reads record k of its 64 lanes from contiguous memory.  The kind,
// slot and layer mask of record k are the same for every lane of a group and
// travel in the kernel arguments.
struct DynRec {
  double A, p1, p2, p3;
  int32_t tidx;  // table base (CMTJ_TABLE)
  int32_t pad;
};

// A record as the kernels hold it.  Where every record is a trapezoid
// (F_KTRAP: no table index to keep) the copy carries the ramp slope A / p2
// in the 8 bytes of tidx and pad, so the division runs once per run instead
// of at every stage time, with the same result.
template <uint32_t FM>
__device__ __forceinline__ DynRec load_rec(const DynRec* p) {
  DynRec r = *p;
  if constexpr ((FM & F_KTRAP) != 0) {
    const double slope = r.A / r.p2;
    __builtin_memcpy(&r.tidx, &slope, sizeof slope);
  }
  return r;
}

// Kernel arguments of one group.  A group shares its structure (layer count,
// per-layer modes, driver kinds and slots, log layout): branches on these
// are wave-uniform.  Numbers differ per lane.
struct GroupArgs {
  const cmtj_junction* js;
  const cmtj_layer* layers;
  const uint32_t* order;  // lane -> junction index
  const DynRec* dyn;
  const double* tables;
  void* log;
  double* mag_out;  // [n][CMTJ_MAX_LAYERS][3] in junction order
  cmtj_status* status;
  cmtj_state* state;  // resumable state [n] in junction order, or null (run from the start)
  const cmtj_stack* stack;  // coupled stacks (k_junction<..., STK>): [stack_n]
  uint32_t stack_n;         // stacks of the launch; member j of stack s is instance s * members + j
  uint32_t stack_pad;       // lanes per stack: a power of two >= members, <= 64 (a stack never straddles a wave)
  const cmtj_group* group;  // dipole-coupled groups (k_junction<..., GRP>): group g on workgroup g
  const double* pair;       // CMTJ_INTERACTION_TENSOR terms [group][j][12][e], or null
  uint32_t group_n;
  uint32_t slice;     // steps per launch when state is set
  uint32_t n;       // lanes of this group
  uint32_t flags;   // feature bits of the group
  int32_t log_f32;
  int32_t nl;       // layers
  int32_t log_mask;
  int32_t mr_mode;
  int32_t ndyn;     // time-dependent driver records
  int32_t pad;
  int32_t lbits[CMTJ_MAX_LAYERS];  // ref_type | mode << 2 | dyn_sot << 4 | alt_stt << 5 | temp << 6
  int32_t dmeta[MAX_DYN];          // kind | slot << 8 | layer mask << 16
  cmtj_constants c;
};

// Driver values of one layer at one time.
template <class R>
struct DV {
  V3<R> H;     // Hext + Hoe + Hdmi + Hres
  V3<R> Hext;  // kept apart for the energy channels
  R K, K2, Jt, Jb, J2t, J2b, I, FL, DL, FL2, DL2, T;
  V3<R> Dt, Db;
};

template <class R>
__device__ __forceinline__ void dv_zero(DV<R>& d) {
  d.H = mk<R>(0, 0, 0);
  d.Hext = d.H;
  d.Dt = d.H;
  d.Db = d.H;
  d.K = d.K2 = d.Jt = d.Jb = d.J2t = d.J2b = d.I = d.FL = d.DL = d.FL2 = d.DL2 = d.T = R(0);
}

// Adds a driver value to the field of its slot.  The slot is wave-uniform, so
// every condition below is a scalar select; adds into fields the kernel never
// reads are removed by the compiler.
template <class R, bool LOGE>
__device__ __forceinline__ void add_slot(DV<R>& d, int slot, R v) {
  const bool isH = slot < 12;  // Hext, Hoe, Hdmi, Hres
  const int axis = slot % 3;
  d.H.x += (isH && axis == 0) ? v : R(0);
  d.H.y += (isH && axis == 1) ? v : R(0);
  d.H.z += (isH && axis == 2) ? v : R(0);
  if (LOGE) {
    d.Hext.x += (slot == CMTJ_SLOT_HEXT + 0) ? v : R(0);
    d.Hext.y += (slot == CMTJ_SLOT_HEXT + 1) ? v : R(0);
    d.Hext.z += (slot == CMTJ_SLOT_HEXT + 2) ? v : R(0);
  }
  d.K += (slot == CMTJ_SLOT_K) ? v : R(0);
  d.K2 += (slot == CMTJ_SLOT_K2) ? v : R(0);
  d.Jt += (slot == CMTJ_SLOT_J_TOP) ? v : R(0);
  d.Jb += (slot == CMTJ_SLOT_J_BOTTOM) ? v : R(0);
  d.J2t += (slot == CMTJ_SLOT_J2_TOP) ? v : R(0);
  d.J2b += (slot == CMTJ_SLOT_J2_BOTTOM) ? v : R(0);
  d.I += (slot == CMTJ_SLOT_CURRENT) ? v : R(0);
  d.FL += (slot == CMTJ_SLOT_FL) ? v : R(0);
  d.DL += (slot == CMTJ_SLOT_DL) ? v : R(0);
  d.FL2 += (slot == CMTJ_SLOT_FL2) ? v : R(0);
  d.DL2 += (slot == CMTJ_SLOT_DL2) ? v : R(0);
  d.T += (slot == CMTJ_SLOT_TEMPERATURE) ? v : R(0);
  const bool isDt = slot >= CMTJ_SLOT_IDMI_TOP && slot < CMTJ_SLOT_IDMI_BOTTOM;
  const bool isDb = slot >= CMTJ_SLOT_IDMI_BOTTOM && slot < CMTJ_NSLOT;
  d.Dt.x += (isDt && axis == 0) ? v : R(0);
  d.Dt.y += (isDt && axis == 1) ? v : R(0);
  d.Dt.z += (isDt && axis == 2) ? v : R(0);
  d.Db.x += (isDb && axis == 0) ? v : R(0);
  d.Db.y += (isDb && axis == 1) ? v : R(0);
  d.Db.z += (isDb && axis == 2) ? v : R(0);
}

// Derived per-layer parameters (in the stage type R).
template <class R>
struct LP {
  R ms_mu0;   // Ms / mu0
  R inv_ms;   // 1 / Ms
  R inv_mst;  // 1 / (Ms thickness)
  R alpha;
  R neg_gamma;
  R conv;     // 1 / (1 + alpha^2)
  R gconv;    // -gamma / (1 + alpha^2)
  V3<R> anis;
  R N[9];
  R Nb[9], Nt[9];  // dipole tensors (bottom, top)
  V3<R> ref, ref2;
  R stt_pref;  // gamma hbar / (e Ms t)
  R P, lam, lam2, kappa, beta, fl_t, dl_t;
  R sig2;  // 2 alpha kB / (Ms V gamma)
};

// Values of the last field/torque evaluation (energy channels).
template <class R>
struct ERec {
  R K, K2, Jb, Jt, I;
  V3<R> Hext, Hiec, Hanis, Hdemag, Hfl, Hdl;
};

#define CMTJ_HAS(b) ((FM & (b)) && ((FM & F_EXACT) || (fl & (b))))

// Quantities of the binary64 anchor evaluation of one layer at (m0, t),
// kept in binary32 for the increment stages of the anchored RK4.
struct Anchor {
  V3<float> m0, H0, P0, Q0;  // m0, total field, m0 x H0, m0 x (m0 x H0)
  float u0;                  // anis . m0
  V3<float> F0, G0;          // STT: m0 x p, m0 x (m0 x p)
  float D0, eta0, stt0;      // STT: eta denominator, eta, gamma aJ eta
  V3<float> C0, CC0, C20, CC20;  // SOT: m0 x p, m0 x (m0 x p), secondary
};

template <class R>
__device__ __forceinline__ V3<R> tensor3(const R* T, V3<R> m) {
  return mk<R>(T[0] * m.x + T[1] * m.y + T[2] * m.z, T[3] * m.x + T[4] * m.y + T[5] * m.z,
               T[6] * m.x + T[7] * m.y + T[8] * m.z);
}

// dm/dt of one layer: effective field (calculateHeffDipoleInjection,
// core/junction.hpp:590-622) and solveLLG (:716-815).
template <class R, uint32_t FM, bool LOGE, bool ANCH = false>
__device__ __forceinline__ V3<R> rhs(const LP<R>& p, int bits, uint32_t fl, V3<R> m,
                                     const DV<R>& d, V3<R> bot, V3<R> top, ERec<R>& er,
                                     Anchor* an = nullptr) {
  V3<R> Nm;
  if (CMTJ_HAS(F_DEMAG_OFF))
    Nm = tensor3(p.N, m);
  else
    Nm = mk<R>(p.N[0] * m.x, p.N[4] * m.y, p.N[8] * m.z);
  const V3<R> Hdemag = Nm * p.ms_mu0;
  V3<R> H = d.H;
  if (ANCH) {
    an->m0 = cvt<float>(m);
    an->u0 = (float)dot(p.anis, m);
  }
  if (CMTJ_HAS(F_ANIS)) {
    const V3<R> Ha = p.anis * ((R(2) * d.K) * dot(p.anis, m) * p.inv_ms);
    H = H + Ha;
    if (LOGE) er.Hanis = Ha;
  } else if (LOGE) {
    er.Hanis = mk<R>(0, 0, 0);
  }
  if (CMTJ_HAS(F_ANIS2)) {
    const R am = dot(p.anis, m);
    H = H + p.anis * ((R(4) * d.K2) * (am * am * am) * p.inv_ms);
  }
  if (CMTJ_HAS(F_IEC) || CMTJ_HAS(F_IEC2)) {
    R sb = d.Jb, st = d.Jt;
    if (CMTJ_HAS(F_IEC2)) {
      sb += R(2) * d.J2b * dot(bot, m);
      st += R(2) * d.J2t * dot(top, m);
    }
    const V3<R> Hiec = (bot * sb + top * st) * p.inv_mst;
    H = H + Hiec;
    if (LOGE) er.Hiec = Hiec;
  } else if (LOGE) {
    er.Hiec = mk<R>(0, 0, 0);
  }
  if (CMTJ_HAS(F_IDMI)) H = H - (cross(d.Db, bot) + cross(d.Dt, top)) * p.inv_mst;
  H = H - Hdemag;
  if (CMTJ_HAS(F_DIPOLE)) H = H - (tensor3(p.Nb, bot) + tensor3(p.Nt, top)) * p.ms_mu0;
  if (LOGE) {
    er.K = d.K;
    er.K2 = d.K2;
    er.Jb = d.Jb;
    er.Jt = d.Jt;
    er.Hext = d.Hext;
    er.Hdemag = Hdemag;
    er.I = R(0);
    er.Hfl = mk<R>(0, 0, 0);
    er.Hdl = er.Hfl;
  }
  const V3<R> prod = cross(m, H);
  const V3<R> prod2 = cross(m, prod);
  const V3<R> dmdt = prod + prod2 * p.alpha;
  if (ANCH) {
    an->H0 = cvt<float>(H);
    an->P0 = cvt<float>(prod);
    an->Q0 = cvt<float>(prod2);
  }
  V3<R> ref = mk<R>(0, 0, 0);
  if (CMTJ_HAS(F_STT) || CMTJ_HAS(F_SOT)) {
    const int rt = bits & 3;
    if (rt == CMTJ_REF_FIXED)
      ref = p.ref;
    else if (rt == CMTJ_REF_TOP)
      ref = top;
    else if (rt == CMTJ_REF_BOTTOM)
      ref = bot;
  }
  if (CMTJ_HAS(F_STT) && ((bits >> 2) & 3) == 1) {
    const R mp = dot(m, ref);
    R eta, den;
    if ((bits >> 5) & 1) {
      den = R(1) + p.lam * mp;
      eta = p.P / den;
    } else {
      den = p.lam2 + R(1) + (p.lam2 - R(1)) * mp;
      eta = (p.P * p.lam2) / den;
    }
    const R stt = p.stt_pref * d.I * eta;
    const V3<R> fieldLike = cross(m, ref);
    const V3<R> dampingLike = cross(m, fieldLike);
    if (ANCH) {
      an->D0 = (float)den;
      an->eta0 = (float)eta;
      an->stt0 = (float)stt;
      an->F0 = cvt<float>(fieldLike);
      an->G0 = cvt<float>(dampingLike);
    }
    if (LOGE) er.I = d.I;
    return (dmdt * p.neg_gamma + dampingLike * (-stt * p.kappa) + fieldLike * (stt * p.beta)) *
           p.conv;
  }
  if (CMTJ_HAS(F_SOT) && ((bits >> 2) & 3) == 2) {
    R Hdl, Hfl, Hdl2 = R(0), Hfl2 = R(0);
    if ((bits >> 4) & 1) {
      Hdl = d.DL;
      Hfl = d.FL;
      Hdl2 = d.DL2;
      Hfl2 = d.FL2;
    } else {
      Hdl = p.dl_t * d.I;
      Hfl = p.fl_t * d.I;
      if (LOGE) er.I = d.I;
    }
    const R a1 = Hfl - p.alpha * Hdl;
    const R b1 = Hdl + p.alpha * Hfl;
    const V3<R> cm = cross(m, ref);
    const V3<R> ccm = cross(m, cm);
    V3<R> r = dmdt + cm * a1 + ccm * b1;
    if (ANCH) {
      an->C0 = cvt<float>(cm);
      an->CC0 = cvt<float>(ccm);
    }
    if (CMTJ_HAS(F_SOT2)) {
      const V3<R> cm2 = cross(m, p.ref2);
      const V3<R> ccm2 = cross(m, cm2);
      r = r + cm2 * (Hfl2 - p.alpha * Hdl2) + ccm2 * (Hdl2 + p.alpha * Hfl2);
      if (ANCH) {
        an->C20 = cvt<float>(cm2);
        an->CC20 = cvt<float>(ccm2);
      }
    }
    if (LOGE) {
      er.Hfl = ref * a1;
      er.Hdl = ref * b1;
    }
    return r * p.gconv;
  }
  return dmdt * p.gconv;
}

// stochasticTorque, core/junction.hpp:824-831.
template <class R>
__device__ __forceinline__ V3<R> stoch(const LP<R>& p, V3<R> m, V3<R> dW) {
  const V3<R> c1 = cross(m, dW);
  const V3<R> c2 = cross(m, c1);
  return (c1 + c2 * p.alpha) * p.gconv;
}

// Increment of dm/dt between the anchor (m0, t) and a stage (m0 + d, t_s):
// f(m0 + d, t_s) - f(m0, t), written so that binary32 rounding acts on
// small quantities only (d, field and torque increments), never on the
// full field or torque.  dd = driver values at t_s minus at t, ds = driver
// values at t_s.  Same terms as rhs(), expanded: products x y become
// dx y_s + x0 dy.
template <uint32_t FM>
__device__ __forceinline__ V3<float> rhs_delta(const LP<float>& p, int bits, uint32_t fl,
                                               const Anchor& A, V3<float> d, const DV<float>& dd,
                                               const DV<float>& ds, V3<float> bot, V3<float> top) {
  V3<float> Nd;
  if (CMTJ_HAS(F_DEMAG_OFF))
    Nd = tensor3(p.N, d);
  else
    Nd = mk<float>(p.N[0] * d.x, p.N[4] * d.y, p.N[8] * d.z);
  V3<float> dH = dd.H - Nd * p.ms_mu0;
  const float du = dot(p.anis, d);
  if (CMTJ_HAS(F_ANIS)) dH = dH + p.anis * (2.f * p.inv_ms * (dd.K * A.u0 + ds.K * du));
  if (CMTJ_HAS(F_ANIS2)) {
    const float us = A.u0 + du;
    const float dcube = du * (us * us + us * A.u0 + A.u0 * A.u0);
    dH = dH + p.anis * (4.f * p.inv_ms * (dd.K2 * (A.u0 * A.u0 * A.u0) + ds.K2 * dcube));
  }
  if (CMTJ_HAS(F_IEC) || CMTJ_HAS(F_IEC2)) {
    float sb = dd.Jb, st = dd.Jt;
    if (CMTJ_HAS(F_IEC2)) {
      sb += 2.f * (dd.J2b * dot(bot, A.m0) + ds.J2b * dot(bot, d));
      st += 2.f * (dd.J2t * dot(top, A.m0) + ds.J2t * dot(top, d));
    }
    dH = dH + (bot * sb + top * st) * p.inv_mst;
  }
  if (CMTJ_HAS(F_IDMI)) dH = dH - (cross(dd.Db, bot) + cross(dd.Dt, top)) * p.inv_mst;
  const V3<float> Hs = A.H0 + dH;
  const V3<float> dP = cross(A.m0, dH) + cross(d, Hs);
  const V3<float> dQ = cross(A.m0, dP) + cross(d, A.P0 + dP);
  const V3<float> ddm = dP + dQ * p.alpha;
  V3<float> ref = mk<float>(0, 0, 0);
  if (CMTJ_HAS(F_STT) || CMTJ_HAS(F_SOT)) {
    const int rt = bits & 3;
    if (rt == CMTJ_REF_FIXED)
      ref = p.ref;
    else if (rt == CMTJ_REF_TOP)
      ref = top;
    else if (rt == CMTJ_REF_BOTTOM)
      ref = bot;
  }
  if (CMTJ_HAS(F_STT) && ((bits >> 2) & 3) == 1) {
    const float dw = dot(d, ref);
    const float c = ((bits >> 5) & 1) ? p.lam : (p.lam2 - 1.f);
    const float deta = -A.eta0 * (c * dw) / (A.D0 + c * dw);
    const float dstt = p.stt_pref * (dd.I * A.eta0 + ds.I * deta);
    const float stts = A.stt0 + dstt;
    const V3<float> dF = cross(d, ref);
    const V3<float> dG = cross(A.m0, dF) + cross(d, A.F0 + dF);
    return (ddm * p.neg_gamma - (dG * stts + A.G0 * dstt) * p.kappa + (dF * stts + A.F0 * dstt) * p.beta) *
           p.conv;
  }
  if (CMTJ_HAS(F_SOT) && ((bits >> 2) & 3) == 2) {
    float Hdl, Hfl, dHdl, dHfl, Hdl2 = 0.f, Hfl2 = 0.f, dHdl2 = 0.f, dHfl2 = 0.f;
    if ((bits >> 4) & 1) {
      Hdl = ds.DL;
      Hfl = ds.FL;
      dHdl = dd.DL;
      dHfl = dd.FL;
      Hdl2 = ds.DL2;
      Hfl2 = ds.FL2;
      dHdl2 = dd.DL2;
      dHfl2 = dd.FL2;
    } else {
      Hdl = p.dl_t * ds.I;
      Hfl = p.fl_t * ds.I;
      dHdl = p.dl_t * dd.I;
      dHfl = p.fl_t * dd.I;
    }
    const V3<float> dC = cross(d, ref);
    const V3<float> dCC = cross(A.m0, dC) + cross(d, A.C0 + dC);
    V3<float> r = ddm + dC * (Hfl - p.alpha * Hdl) + A.C0 * (dHfl - p.alpha * dHdl) +
                  dCC * (Hdl + p.alpha * Hfl) + A.CC0 * (dHdl + p.alpha * dHfl);
    if (CMTJ_HAS(F_SOT2)) {
      const V3<float> dC2 = cross(d, p.ref2);
      const V3<float> dCC2 = cross(A.m0, dC2) + cross(d, A.C20 + dC2);
      r = r + dC2 * (Hfl2 - p.alpha * Hdl2) + A.C20 * (dHfl2 - p.alpha * dHdl2) +
          dCC2 * (Hdl2 + p.alpha * Hfl2) + A.CC20 * (dHdl2 + p.alpha * dHfl2);
    }
    return r * p.gconv;
  }
  return ddm * p.gconv;
}

template <class R>
__device__ __forceinline__ void load_layer(const cmtj_layer& L, const cmtj_constants& c, LP<R>& p,
                                           DV<R>& base, V3<double>& m) {
  p.ms_mu0 = (R)(L.Ms / c.mu0);
  p.inv_ms = (R)(1.0 / L.Ms);
  p.inv_mst = (R)(1.0 / (L.Ms * L.thickness));
  p.alpha = (R)L.damping;
  p.neg_gamma = (R)(-c.gyro);
  p.conv = (R)(1.0 / (1.0 + L.damping * L.damping));
  p.gconv = (R)(-c.gyro / (1.0 + L.damping * L.damping));
  p.anis = mk<R>((R)L.anis[0], (R)L.anis[1], (R)L.anis[2]);
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    p.N[q] = (R)L.demag[q];
    p.Nb[q] = (R)L.dipole_bottom[q];
    p.Nt[q] = (R)L.dipole_top[q];
  }
  p.ref = mk<R>((R)L.ref[0], (R)L.ref[1], (R)L.ref[2]);
  p.ref2 = mk<R>((R)L.ref2[0], (R)L.ref2[1], (R)L.ref2[2]);
  p.stt_pref = (R)(c.gyro * c.hbar / (c.echarge * L.Ms * L.thickness));
  p.P = (R)L.polarisation;
  p.lam = (R)L.slonczewski;
  p.lam2 = (R)(L.slonczewski * L.slonczewski);
  p.kappa = (R)L.kappa;
  p.beta = (R)L.beta;
  p.fl_t = (R)L.fl_torque;
  p.dl_t = (R)L.dl_torque;
  const double V = L.cell_surface * L.thickness;
  p.sig2 = (R)((2.0 * L.damping * c.kB) / (L.Ms * V * c.gyro));
  dv_zero(base);
  for (int s = 0; s < CMTJ_NSLOT; ++s) {
    const int ty = L.drv[s].type;
    const double cv = (ty == CMTJ_TABLE) ? 0.0 : L.drv[s].constant;
    if (cv != 0.0) add_slot<R, true>(base, s, (R)cv);
  }
  m = mk<double>(L.mag[0], L.mag[1], L.mag[2]);
}

template <class A, class B>
__device__ __forceinline__ DV<A> dv_cvt(const DV<B>& s) {
  DV<A> d;
  d.H = cvt<A>(s.H);
  d.Hext = cvt<A>(s.Hext);
  d.Dt = cvt<A>(s.Dt);
  d.Db = cvt<A>(s.Db);
  d.K = (A)s.K;
  d.K2 = (A)s.K2;
  d.Jt = (A)s.Jt;
  d.Jb = (A)s.Jb;
  d.J2t = (A)s.J2t;
  d.J2b = (A)s.J2b;
  d.I = (A)s.I;
  d.FL = (A)s.FL;
  d.DL = (A)s.DL;
  d.FL2 = (A)s.FL2;
  d.DL2 = (A)s.DL2;
  d.T = (A)s.T;
  return d;
}

// (a - b) in binary64, rounded once to binary32
__device__ __forceinline__ DV<float> dv_diff(const DV<double>& a, const DV<double>& b) {
  DV<float> d;
  d.H = cvt<float>(a.H - b.H);
  d.Hext = cvt<float>(a.Hext - b.Hext);
  d.Dt = cvt<float>(a.Dt - b.Dt);
  d.Db = cvt<float>(a.Db - b.Db);
  d.K = (float)(a.K - b.K);
  d.K2 = (float)(a.K2 - b.K2);
  d.Jt = (float)(a.Jt - b.Jt);
  d.Jb = (float)(a.Jb - b.Jb);
  d.J2t = (float)(a.J2t - b.J2t);
  d.J2b = (float)(a.J2b - b.J2b);
  d.I = (float)(a.I - b.I);
  d.FL = (float)(a.FL - b.FL);
  d.DL = (float)(a.DL - b.DL);
  d.FL2 = (float)(a.FL2 - b.FL2);
  d.DL2 = (float)(a.DL2 - b.DL2);
  d.T = (float)(a.T - b.T);
  return d;
}

// Numbers of the first DREG records stay in registers for the whole run.
constexpr int DREG = 2;

// Field routing of the first DREG records: axis[k] is the field axis record
// k feeds (-1 if it feeds no field slot; wave-uniform), on[k][l] whether it
// feeds layer l (uniform for the all-layers-in-a-lane kernels, per lane for
// the one-layer-per-lane ones).  Values are routed by selects, never by
// multiplying with a 0/1 weight: 0 * inf would turn a non-finite drive of
// one layer into a NaN of its neighbours, which the reference never sees.
template <class R, int LM>
struct FieldWeights {
  int axis[DREG];
  bool on[DREG][LM];
  double win_lo, win_hi;  // step-driver kernels: every record is zero outside [win_lo, win_hi]
};

template <class R, int LM>
__device__ __forceinline__ void field_weights(FieldWeights<R, LM>& fw, const GroupArgs& g, int lsh = 0,
                                              const DynRec* dr = nullptr) {
  fw.win_lo = __builtin_inf();
  fw.win_hi = -__builtin_inf();
  if (dr) {
#pragma unroll
    for (int k = 0; k < DREG; ++k) {
      if (k < g.ndyn) {
        fw.win_lo = fmin(fw.win_lo, dr[k].p1);
        fw.win_hi = fmax(fw.win_hi, dr[k].p2);
      }
    }
  }
#pragma unroll
  for (int k = 0; k < DREG; ++k) {
    const int meta = (k < g.ndyn) ? g.dmeta[k] : 0;
    const int type = meta & 0xff;
    const int slot = (meta >> 8) & 0xff;
    const int lmask = (meta >> 16) & 0xff;
    const bool isH = type != CMTJ_CONSTANT && slot < CMTJ_SLOT_K;
    fw.axis[k] = isH ? slot % 3 : -1;
#pragma unroll
    for (int l = 0; l < LM; ++l) fw.on[k][l] = isH && ((lmask >> (l + lsh)) & 1);
  }
}

// H[axis] += v where the record feeds this layer.  The layer choice is a
// select (another layer never sees v); the axis (wave-uniform) is a 0/1
// weight, so a non-finite v reaches only the layer it drives.
template <class R>
__device__ __forceinline__ void add_field(V3<R>& H, int axis, bool on, R v) {
  const R vv = on ? v : R(0);
  H.x = fma(R(axis == 0 ? 1 : 0), vv, H.x);
  H.y = fma(R(axis == 1 ? 1 : 0), vv, H.y);
  H.z = fma(R(axis == 2 ? 1 : 0), vv, H.z);
}

template <uint32_t FM>
__device__ __forceinline__ DynRec pick_rec(const DynRec (&dr)[DREG], const GroupArgs& g, int k, uint32_t lane) {
  DynRec r;
  if (k < DREG) {
#pragma unroll
    for (int q = 0; q < DREG; ++q)
      if (q == k) r = dr[q];
  } else {
    r = load_rec<FM>(g.dyn + (size_t)k * g.n + lane);
  }
  return r;
}

__device__ __forceinline__ bool transcendental(int type) {
  return type == CMTJ_SINE || type == CMTJ_POSINE || type == CMTJ_HALFSINE || type == CMTJ_GAUSSIMPULSE ||
         type == CMTJ_GAUSSSTEP;
}

// Value of one time-dependent record at t.  FT: binary32 transcendentals
// (phase and exponent argument still formed in binary64) for the anchored
// FP32 mode, whose stage increments need the drive only to ~1e-7 relative.
template <class R, bool FT = false, uint32_t FM = 0>
__device__ __forceinline__ R wave_value(int type, const DynRec& r, const GroupArgs& g, double t, uint32_t step,
                                        int stage, int sps) {
  if constexpr ((FM & F_KSTEP) != 0) {  // stepUpdate (drivers.hpp:73-78), the only kind of this group
    return (t >= r.p1 && t <= r.p2) ? (R)r.A : R(0);
  } else if constexpr ((FM & F_KSINE) != 0) {
    if (sizeof(R) == 4 || FT) return (R)wave_part_f_trans(CMTJ_SINE, r.A, r.p1, r.p2, t);
    return (R)wave_part_trans(CMTJ_SINE, r.A, r.p1, r.p2, t);
  } else if constexpr ((FM & F_KTRAP) != 0) {  // the only kind of this group (slope from load_rec)
    double slope;
    __builtin_memcpy(&slope, &r.tidx, sizeof slope);
    return (R)wave_trapezoid_slope(r.A, slope, r.p1, r.p2, r.p3, t);
  }
  if (type == CMTJ_TABLE) return (R)g.tables[(size_t)r.tidx + (size_t)step * sps + stage];
  if (sizeof(R) == 4) return (R)wave_part_f(type, r.A, r.p1, r.p2, r.p3, t);
  if (FT && transcendental(type)) return (R)wave_part_f_trans(type, r.A, r.p1, r.p2, t);
  return (R)wave_part(type, r.A, r.p1, r.p2, r.p3, t);
}

// v0[k] = v for the first DREG records (k wave-uniform, possibly a runtime index)
__device__ __forceinline__ void keep_v0(double* v0, int k, double v) {
  if (!v0) return;
#pragma unroll
  for (int q = 0; q < DREG; ++q)
    if (q == k) v0[q] = v;
}

// Driver values at time t: the per-layer base (constant parts) plus every
// time-dependent record.  Kinds, slots and layer masks are wave-uniform.
// Records that feed only field slots take the fast path (a select and an
// add on the uniform axis); other slots go through add_slot.  v0 (optional)
// receives the values of the first DREG records.
template <class R, int LM, uint32_t FM, bool LOGE, bool FT = false>
__device__ __forceinline__ void eval_drivers(DV<R> (&dv)[LM], const DV<R> (&base)[LM],
                                             const DynRec (&dr)[DREG], const FieldWeights<R, LM>& fw,
                                             const GroupArgs& g, uint32_t lane, uint32_t fl, double t,
                                             uint32_t step, int stage, int sps, int lsh = 0,
                                             double* v0 = nullptr) {
#pragma unroll
  for (int l = 0; l < LM; ++l) dv[l] = base[l];
  if constexpr ((FM & F_KSTEP) != 0) {  // stepUpdate is zero outside [t_start, t_stop]
    if (!(t >= fw.win_lo && t <= fw.win_hi)) {
      if (v0)
#pragma unroll
        for (int k = 0; k < DREG; ++k) v0[k] = 0.0;
      return;
    }
  }
  if (CMTJ_HAS(F_DYN) || (LOGE && CMTJ_HAS(F_DYNH))) {
    for (int k = 0; k < g.ndyn; ++k) {
      const int meta = g.dmeta[k];
      const int type = meta & 0xff;
      const int slot = (meta >> 8) & 0xff;
      const int lmask = (meta >> 16) & 0xff;
      const DynRec r = pick_rec<FM>(dr, g, k, lane);
      const R v = wave_value<R, FT, FM>(type, r, g, t, step, stage, sps);
      keep_v0(v0, k, (double)v);
#pragma unroll
      for (int l = 0; l < LM; ++l)
        if ((lmask >> (l + lsh)) & 1) add_slot<R, LOGE>(dv[l], slot, v);
    }
  } else if (CMTJ_HAS(F_DYNH)) {
#pragma unroll
    for (int k = 0; k < DREG; ++k) {
      if (k < g.ndyn) {
        const int type = g.dmeta[k] & 0xff;
        const R v = wave_value<R, FT, FM>(type, dr[k], g, t, step, stage, sps);
        if (v0) v0[k] = (double)v;
#pragma unroll
        for (int l = 0; l < LM; ++l) add_field(dv[l].H, fw.axis[k], fw.on[k][l], v);
      }
    }
    for (int k = DREG; !(FM & F_EXACT) && k < g.ndyn; ++k) {
      const int meta = g.dmeta[k];
      const DynRec r = pick_rec<FM>(dr, g, k, lane);
      const R v = wave_value<R, FT, FM>(meta & 0xff, r, g, t, step, stage, sps);
#pragma unroll
      for (int l = 0; l < LM; ++l)
        if (((meta >> 16) >> (l + lsh)) & 1) add_slot<R, false>(dv[l], (meta >> 8) & 0xff, v);
    }
  }
}

// Record k's value at t0: from v0 for the first DREG records, else recomputed.
template <uint32_t FM>
__device__ __forceinline__ double value_t0(const double (&v0)[DREG], int k, int type, const DynRec& r,
                                           const GroupArgs& g, double t0, uint32_t step, int sps) {
  if (k < DREG) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < DREG; ++q)
      if (q == k) v = v0[q];
    return v;
  }
  return wave_value<double, true, FM>(type, r, g, t0, step, 0, sps);
}

// Driver increments dd = dv(ts) - dv(t0) of the time-dependent records of
// the anchored FP32 mode (binary64 difference, one rounding to binary32;
// constant parts cancel).  v0: the record values at t0 from eval_drivers.
template <int LM, uint32_t FM>
__device__ __forceinline__ void eval_driver_deltas(DV<float> (&dd)[LM], const DynRec (&dr)[DREG],
                                                   const FieldWeights<float, LM>& fw, const GroupArgs& g,
                                                   uint32_t lane, uint32_t fl, double t0, double ts,
                                                   uint32_t step, int stage, int sps, const double (&v0)[DREG],
                                                   int lsh = 0) {
#pragma unroll
  for (int l = 0; l < LM; ++l) dv_zero(dd[l]);
  if constexpr ((FM & F_KSTEP) != 0) {  // both times outside every step window: no increment
    if (!(ts >= fw.win_lo && t0 <= fw.win_hi)) return;
  }
  if (CMTJ_HAS(F_DYN)) {
    for (int k = 0; k < g.ndyn; ++k) {
      const int meta = g.dmeta[k];
      const int type = meta & 0xff;
      const DynRec r = pick_rec<FM>(dr, g, k, lane);
      const double v = wave_value<double, true, FM>(type, r, g, ts, step, stage, sps) -
                       value_t0<FM>(v0, k, type, r, g, t0, step, sps);
#pragma unroll
      for (int l = 0; l < LM; ++l)
        if (((meta >> 16) >> (l + lsh)) & 1) add_slot<float, false>(dd[l], (meta >> 8) & 0xff, (float)v);
    }
  } else if (CMTJ_HAS(F_DYNH)) {
#pragma unroll
    for (int k = 0; k < DREG; ++k) {
      if (k < g.ndyn) {
        const int type = g.dmeta[k] & 0xff;
        const float v = (float)(wave_value<double, true, FM>(type, dr[k], g, ts, step, stage, sps) - v0[k]);
#pragma unroll
        for (int l = 0; l < LM; ++l) add_field(dd[l].H, fw.axis[k], fw.on[k][l], v);
      }
    }
    for (int k = DREG; !(FM & F_EXACT) && k < g.ndyn; ++k) {
      const int meta = g.dmeta[k];
      const DynRec r = pick_rec<FM>(dr, g, k, lane);
      const double v = wave_value<double, true, FM>(meta & 0xff, r, g, ts, step, stage, sps) -
                       wave_value<double, true, FM>(meta & 0xff, r, g, t0, step, 0, sps);
#pragma unroll
      for (int l = 0; l < LM; ++l)
        if (((meta >> 16) >> (l + lsh)) & 1) add_slot<float, false>(dd[l], (meta >> 8) & 0xff, (float)v);
    }
  }
}

template <class R>
__device__ __forceinline__ DV<R> dv_add(const DV<R>& a, const DV<R>& b) {
  DV<R> d;
  d.H = a.H + b.H;
  d.Hext = a.Hext + b.Hext;
  d.Dt = a.Dt + b.Dt;
  d.Db = a.Db + b.Db;
  d.K = a.K + b.K;
  d.K2 = a.K2 + b.K2;
  d.Jt = a.Jt + b.Jt;
  d.Jb = a.Jb + b.Jb;
  d.J2t = a.J2t + b.J2t;
  d.J2b = a.J2b + b.J2b;
  d.I = a.I + b.I;
  d.FL = a.FL + b.FL;
  d.DL = a.DL + b.DL;
  d.FL2 = a.FL2 + b.FL2;
  d.DL2 = a.DL2 + b.DL2;
  d.T = a.T + b.T;
  return d;
}

// CVector::normalize (core/cvector.hpp:185-192): divide by the length.
__device__ __forceinline__ V3<double> normalize_exact(V3<double> a) {
  const double mag = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
  if (mag != 0.0) {
    a.x = a.x / mag;
    a.y = a.y / mag;
    a.z = a.z / mag;
  }
  return a;
}

// Same result to ~1 ulp: binary32 reciprocal square root refined by two
// Newton steps in binary64 (used by the FP32-bulk kernels).
__device__ __forceinline__ V3<double> normalize_fast(V3<double> a) {
  const double s = a.x * a.x + a.y * a.y + a.z * a.z;
  if (s == 0.0) return a;
  double y = (double)__frsqrt_rn((float)s);
  y = y * (1.5 - 0.5 * s * y * y);
  y = y * (1.5 - 0.5 * s * y * y);
  return mk<double>(a.x * y, a.y * y, a.z * y);
}

template <class R>
__device__ __forceinline__ V3<double> normalize_m(V3<double> a) {
  if (sizeof(R) == 8) return normalize_exact(a);
  return normalize_fast(a);
}

// Channels of one junction's log entry (logLayerParams, core/junction.hpp:
// 1580-1636): m, resistance, energies.
__device__ __forceinline__ int log_channels(int nl, int mr, int mask) {
  int n = 0;
  if (mask & CMTJ_LOG_MX) n += nl;
  if (mask & CMTJ_LOG_MYZ) n += 2 * nl;
  if (mask & CMTJ_LOG_MR) n += (mr == CMTJ_MR_CLASSIC) ? 1 : (mr == CMTJ_MR_STRIP ? 2 : 0);
  if (mask & CMTJ_LOG_ENERGY) n += CMTJ_ENERGY_CHANNELS * nl;
  if (mask & CMTJ_LOG_TIME) n += 1;
  return n;
}

// One log value at byte address p (binary32 or binary64 log).
__device__ __forceinline__ void log_put(bool f32, char* p, double v) {
  if (f32)
    *(float*)p = (float)v;
  else
    *(double*)p = v;
}

// Resistance coefficients of one layer, read from the descriptors once per
// run: the descriptors may sit in host-mapped memory (a serving loop's
// plans), where every read crosses the host link.  ahe2 = 0.5 AHE and
// xy = AMR_Y + SMR_Y are the first products of the reference's expressions
// (core/junction.hpp:1840-1846), so the entries round as before.
struct MRCoef {
  double ref[3];  // CLASSIC with one layer: the fixed reference vector
  double rx0, amrx, smrx, ry0, ahe2, xy;  // STRIP
};

__device__ __forceinline__ MRCoef mr_coef(const GroupArgs& g, const cmtj_junction* J, int l, bool on) {
  MRCoef c = {{0.0, 0.0, 0.0}, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (!on || !(g.log_mask & CMTJ_LOG_MR)) return c;
  if (g.mr_mode == CMTJ_MR_CLASSIC && g.nl == 1) {
    const double* rf = g.layers[J->layer_offset].ref;
    c.ref[0] = rf[0];
    c.ref[1] = rf[1];
    c.ref[2] = rf[2];
  } else if (g.mr_mode == CMTJ_MR_STRIP) {
    c.rx0 = J->Rx0[l];
    c.amrx = J->AMR_X[l];
    c.smrx = J->SMR_X[l];
    c.ry0 = J->Ry0[l];
    c.ahe2 = 0.5 * J->AHE[l];
    c.xy = J->AMR_Y[l] + J->SMR_Y[l];
  }
  return c;
}

// Layers of capacity whose STRIP coefficients a one-lane-per-instance kernel
// keeps in registers; larger kernels read the descriptors per entry.
constexpr int kMrRegLayers = 2;

// logLayerParams (core/junction.hpp:1580-1636) for entry k.  p points at
// channel 0 of entry k; consecutive channels are cs bytes apart.  mc: the
// resistance coefficients of layers 0..min(LM, kMrRegLayers)-1.
template <class R, int LM, uint32_t FM, bool LOGE>
__device__ __forceinline__ void write_entry(const GroupArgs& g, const cmtj_junction* J,
                                            const V3<double> (&m)[LM], const ERec<R> (&er)[LM], char* p,
                                            uint64_t cs, double Rp, double Rap,
                                            const MRCoef (&mc)[LM < kMrRegLayers ? LM : kMrRegLayers]) {
  const int nl = g.nl, lmask = g.log_mask, mr = g.mr_mode;
  const bool f32 = g.log_f32;
#pragma unroll
  for (int l = 0; l < LM; ++l) {
    if (l < nl) {
      if (lmask & CMTJ_LOG_MX) {
        log_put(f32, p, m[l].x);
        p += cs;
      }
      if (lmask & CMTJ_LOG_MYZ) {
        log_put(f32, p, m[l].y);
        p += cs;
        log_put(f32, p, m[l].z);
        p += cs;
      }
    }
  }
  if (lmask & CMTJ_LOG_MR) {
    if (mr == CMTJ_MR_CLASSIC) {
      // calculateMagnetoresistance (:1865-1867); 1 layer against the fixed
      // reference vector, 2 layers against each other (:1623-1629)
      double cosT;
      if (nl == 1) {
        const double* rf = mc[0].ref;
        cosT = m[0].x * rf[0] + m[0].y * rf[1] + m[0].z * rf[2];
      } else {
        const V3<double>& m1 = m[LM > 1 ? 1 : 0];
        cosT = m[0].x * m1.x + m[0].y * m1.y + m[0].z * m1.z;
      }
      log_put(f32, p, Rp + (((Rap - Rp) / 2.0) * (1.0 - cosT)));
      p += cs;
    } else if (mr == CMTJ_MR_STRIP) {
      // stripMagnetoResistance (:1835-1857), parallel sum of the layers
      double rx = 0.0, ry = 0.0;
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          MRCoef q;
          if (l < kMrRegLayers)
            q = mc[l < kMrRegLayers ? l : 0];
          else
            q = mr_coef(g, J, l, true);
          const double Rx = q.rx0 + q.amrx * (m[l].x * m[l].x) + q.smrx * (m[l].y * m[l].y);
          const double Ry = q.ry0 + q.ahe2 * m[l].z + q.xy * m[l].x * m[l].y;
          rx += 1. / Rx;
          ry += 1. / Ry;
        }
      }
      log_put(f32, p, 1 / rx);
      p += cs;
      log_put(f32, p, 1 / ry);
      p += cs;
    }
  }
  if (LOGE && (lmask & CMTJ_LOG_ENERGY)) {
#pragma unroll
    for (int l = 0; l < LM; ++l) {
      if (l < nl) {
        const ERec<R>& e = er[l];
        const double v[CMTJ_ENERGY_CHANNELS] = {
            e.K, e.K2, e.Jb, e.Jt, e.I, e.Hext.x, e.Hext.y, e.Hext.z,
            e.Hiec.x, e.Hiec.y, e.Hiec.z, e.Hanis.x, e.Hanis.y, e.Hanis.z,
            e.Hdemag.x, e.Hdemag.y, e.Hdemag.z, e.Hfl.x, e.Hfl.y, e.Hfl.z,
            e.Hdl.x, e.Hdl.y, e.Hdl.z};
#pragma unroll
        for (int q = 0; q < CMTJ_ENERGY_CHANNELS; ++q) {
          log_put(f32, p, v[q]);
          p += cs;
        }
      }
    }
  }
}

// Current of stack member j for the stage times of the step at t
// (Stack<T>::runSimulationKCL / runSimulationNonKCL, core/stack.hpp:305-410):
// the pair couplings come from the members' magnetisations at the start of
// the step (storeMagnetisations, :65-76) and travel between the lanes of the
// wave.  I[s]: the current at stage time s (RK4: t, t + dt/2, t + dt);
// Ilog, Clog: the stack log's I_j and C_j_{j+1} of this step.
// The stack's constants of one member, read once before the step loop (the
// loop's log stores may alias the descriptor, so loads inside it would be
// repeated every step).
struct StackLane {
  cmtj_driver current;  // the stack current driver
  double pol[3];
  double cj;      // coupling of pair (j, j+1); 0 for the last member
  double phase;   // phase shift of member j (non-KCL)
  int two_layer, series, kcl;
  bool nd;        // non-KCL, non-delayed: the members step one after another
};

__device__ __forceinline__ StackLane stack_lane(const cmtj_stack* sk, int j, int n) {
  StackLane s;
  s.current = sk->current;
  s.pol[0] = sk->pol[0];
  s.pol[1] = sk->pol[1];
  s.pol[2] = sk->pol[2];
  s.cj = (j + 1 < n) ? sk->coupling[j] : 0.0;
  s.two_layer = sk->two_layer;
  s.series = sk->series;
  s.kcl = sk->kcl;
  s.phase = s.series ? sk->phase_offset * j : sk->phase_offset;
  s.nd = !s.kcl && !sk->delayed;
  return s;
}

// m_top . p of a member whose layers are in this lane's registers (p: its
// bottom layer, or the stack's polarisation vector), c_dot of
// core/junction.hpp:82-84.
template <int LM>
__device__ __forceinline__ double stack_d(const StackLane& sk, int ti, int bi, const V3<double> (&m)[LM]) {
  V3<double> mt = m[0], mb = m[0];
#pragma unroll
  for (int l = 0; l < LM; ++l) {
    mt = (l == ti) ? m[l] : mt;
    mb = (l == bi) ? m[l] : mb;
  }
  const V3<double> p = sk.two_layer ? mb : mk<double>(sk.pol[0], sk.pol[1], sk.pol[2]);
  return mt.x * p.x + mt.y * p.y + mt.z * p.z;
}

// d: this member's m_top . p; member q of the stack holds its copy in wave
// lane base + q stride (stride: lanes per member).
__device__ __forceinline__ void stack_current(const StackLane& sk, int j, int n, int base, int stride, double d,
                                              double t, double dt, const double* tab, uint32_t step,
                                              double (&I)[3], double& Ilog, double& Clog) {
  const double dn = __shfl(d, base + ((j + 1 < n) ? j + 1 : j) * stride, 64);
  const double Cj = (j + 1 < n) ? sk.cj * (sk.series ? d + dn : d - dn) : 0.0;
  if (sk.kcl) {
    // totalCurrent = I0 + sum_j C_j I0, the same constant for every member;
    // a callback current is tabulated once per step
    const double I0 = (sk.current.type == CMTJ_TABLE) ? tab[(size_t)sk.current.table_index + step]
                                                      : driver_value(sk.current, t);
    double total = I0;
    // unrolled so the lane reads of several pairs are in flight together;
    // the sum keeps the reference's order
#pragma unroll 4
    for (int q = 0; q + 1 < n; ++q) total += __shfl(Cj, base + q * stride, 64) * I0;
    I[0] = I[1] = I[2] = total;
    Ilog = total;
    Clog = Cj;
  } else {
    // effectiveCoupling *= (1 + C_{j-1}) along the chain; the member's driver
    // is current * effectiveCoupling with its phase shifted
    double eff = 1.0;
#pragma unroll 4
    for (int q = 0; q + 1 < n; ++q) {
      const double cq = __shfl(Cj, base + q * stride, 64);
      if (q < j) eff *= (1 + cq);
    }
    Clog = eff * (1 + Cj);
    cmtj_driver dl = sk.current;
    dl.constant *= eff;
    dl.amplitude *= eff;
    dl.phase += sk.phase;
    I[0] = driver_value(dl, t);
    I[1] = driver_value(dl, t + 0.5 * dt);
    I[2] = driver_value(dl, t + dt);
    Ilog = I[0];
  }
}

// Non-delayed non-KCL coupling (runSimulationNonKCL with delayed == false,
// core/stack.hpp:334-343): within a step the members step in order, and
// member q's coupling to member q-1 is read from q-1's magnetisations after
// its step and q's before it.  Called by every lane of the stack at sub-step
// q = 0..n-1 (member q steps at sub-step q) with the stack's running product
// eff, 1 at q = 0: eff *= 1 + c_{q-1} (d_{q-1} +- d_q), d = m_top . m_bottom.
// Member q's current is the stack driver times eff with its phase shifted;
// the stack log's C_{q-1}_q is eff after the update.
template <int LM>
__device__ __forceinline__ void stack_current_nd(const StackLane& sk, int j, int n, int base, int ti, int bi,
                                                 const V3<double> (&m)[LM], double t, double dt, int q,
                                                 double& eff, double (&I)[3], double& Ilog, double& Clog) {
  V3<double> mt = m[0], mb = m[0];
#pragma unroll
  for (int l = 0; l < LM; ++l) {
    mt = (l == ti) ? m[l] : mt;
    mb = (l == bi) ? m[l] : mb;
  }
  const double d = mt.x * mb.x + mt.y * mb.y + mt.z * mb.z;
  if (q > 0) {
    const double dp = __shfl(d, base + q - 1, 64), dq = __shfl(d, base + q, 64);
    const double cp = __shfl(sk.cj, base + q - 1, 64);  // coupling of pair (q-1, q)
    eff *= (1 + cp * (sk.series ? dp + dq : dp - dq));
    if (j == q - 1) Clog = eff;
  }
  if (j == q) {
    cmtj_driver dl = sk.current;
    dl.constant *= eff;
    dl.amplitude *= eff;
    dl.phase += sk.phase;
    I[0] = driver_value(dl, t);
    I[1] = driver_value(dl, t + 0.5 * dt);
    I[2] = driver_value(dl, t + dt);
    Ilog = I[0];
    if (q == n - 1) Clog = eff;
  }
  (void)n;
}

// A dipole-coupled group in LDS: the top-layer magnetisations of the step,
// positions, source strengths s and the inverse distances il[j][e] = 1/|r_e
// - r_j| (0 on the diagonal, so a lane's own term vanishes without a
// branch).  il is [j][e]: the lanes of one source read consecutive words;
// m, pos and s of a source are broadcast reads.  With GW waves per group,
// waves 1..GW-1 sum their share of the sources into part[w - 1] and ctl
// tells them whether another step follows.  About 38 KB at 64 elements and
// one wave (four groups per CU), 42 KB with four waves.
template <int GW>
struct GroupLds {
  double m[CMTJ_MAX_GROUP][3];
  double pos[CMTJ_MAX_GROUP][3];
  double s[CMTJ_MAX_GROUP];
  double il[CMTJ_MAX_GROUP][CMTJ_MAX_GROUP];
  double part[GW > 1 ? GW - 1 : 1][CMTJ_MAX_GROUP][3];
  int ctl;
};

// Field on element e of the sources j0 <= j < j1, from the top-layer
// magnetisations in LDS (stepFunctionalSolver, core/reservoir.hpp:126-147,
// sources ascending):
//   sum_j s_j (3 (m_j . r) r / |r|^5 - m_j / |r|^3),  r = r_e - r_j,
// which is computeDipoleInteraction (:79-96) with s_j = (Ms_j / mu0) V_j,
// and computeDipoleInteractionNoumra (:106-119) with s_j = Ms_j V_j / mu0 /
// (4 pi): its tensor (r r^T - |r|^2/3) 3 / (4 pi |r|^5) applied to m_j is
// the same point-dipole field over 4 pi.  Powers of the stored 1/|r| replace
// the reference's pow() and divisions (a few ulp apart).
template <int GW>
__device__ __forceinline__ V3<double> group_field(const GroupLds<GW>& L, int j0, int j1, int e) {
  V3<double> H = mk<double>(0, 0, 0);
  const V3<double> ri = mk<double>(L.pos[e][0], L.pos[e][1], L.pos[e][2]);
#pragma unroll 4
  for (int j = j0; j < j1; ++j) {
    const V3<double> r = ri - mk<double>(L.pos[j][0], L.pos[j][1], L.pos[j][2]);
    const V3<double> mj = mk<double>(L.m[j][0], L.m[j][1], L.m[j][2]);
    const double il = L.il[j][e], il2 = il * il;
    const double b = L.s[j] * (il2 * il);                                          // s / |r|^3
    const double a = 3 * (mj.x * r.x + mj.y * r.y + mj.z * r.z) * (b * il2);      // 3 s (m . r) / |r|^5
    H.x = fma(a, r.x, fma(-b, mj.x, H.x));
    H.y = fma(a, r.y, fma(-b, mj.y, H.y));
    H.z = fma(a, r.z, fma(-b, mj.z, H.z));
  }
  return H;
}

// Field on element e of the sources j0 <= j < j1 from per-pair affine
// terms (CMTJ_INTERACTION_TENSOR): f_j = T_ej m_j + b_ej, with the group's
// terms pt laid out [j][12][e] (the lanes of one source and component read
// consecutive words), the sum over j ascending as the reference adds the
// callable's vectors (core/reservoir.hpp:135-142), the diagonal skipped.
template <int GW>
__device__ __forceinline__ V3<double> group_field_tensor(const GroupLds<GW>& L, const double* pt, int n, int j0,
                                                         int j1, int e) {
  V3<double> H = mk<double>(0, 0, 0);
  for (int j = j0; j < j1; ++j) {
    if (j == e) continue;
    const double* p = pt + (size_t)j * 12 * (size_t)n + (size_t)e;
    const double mx = L.m[j][0], my = L.m[j][1], mz = L.m[j][2];
    H.x += fma(p[0], mx, fma(p[(size_t)n], my, fma(p[2 * (size_t)n], mz, p[9 * (size_t)n])));
    H.y += fma(p[3 * (size_t)n], mx, fma(p[4 * (size_t)n], my, fma(p[5 * (size_t)n], mz, p[10 * (size_t)n])));
    H.z += fma(p[6 * (size_t)n], mx, fma(p[7 * (size_t)n], my, fma(p[8 * (size_t)n], mz, p[11 * (size_t)n])));
  }
  return H;
}

// Sources of wave w of a group of n elements on GW waves.
template <int GW>
__device__ __forceinline__ void group_share(int n, int w, int& j0, int& j1) {
  const int chunk = (n + GW - 1) / GW;
  j0 = min(n, w * chunk);
  j1 = min(n, j0 + chunk);
}

// Waves 1..GW-1 of a group: per step, wait for the magnetisations (two
// barriers, as wave 0 writes them between its first two), sum the field of
// this wave's sources on every element, publish it (third barrier).  Wave 0
// sets ctl to 0 between the first two barriers of the round after its last
// step, so every wave passes the same barriers and leaves together.
template <int GW>
__device__ __noinline__ void group_helper(GroupLds<GW>& L, int n, int w, const double* pt) {
  const int e = (int)(threadIdx.x & 63);
  int j0, j1;
  group_share<GW>(n, w, j0, j1);
  __syncthreads();  // positions and strengths are written
  for (;;) {
    __syncthreads();  // the previous step's reads are done
    __syncthreads();  // m and ctl are written
    if (L.ctl == 0) break;
    if (e < n) {
      const V3<double> h = pt ? group_field_tensor<GW>(L, pt, n, j0, j1, e) : group_field<GW>(L, j0, j1, e);
      L.part[w - 1][e][0] = h.x;
      L.part[w - 1][e][1] = h.y;
      L.part[w - 1][e][2] = h.z;
    }
    __syncthreads();  // partial fields are written
  }
}

template <class R, int LM>
__device__ __forceinline__ void add_field_to(DV<R> (&dv)[LM], int layer, V3<double> H) {
#pragma unroll
  for (int l = 0; l < LM; ++l)
    if (l == layer) dv[l].H = dv[l].H + cvt<R>(H);
}

template <class R, int LM>
__device__ __forceinline__ void set_current(DV<R> (&dv)[LM], double I) {
#pragma unroll
  for (int l = 0; l < LM; ++l) dv[l].I = (R)I;
}

template <class R, int LM, int SOLVER, uint32_t FM, bool LOGE, bool STK = false, bool GRP = false, int GW = 1>
__device__ __forceinline__ void junction_body(const GroupArgs& g) {
  uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
  [[maybe_unused]] const cmtj_group* gp = nullptr;
  [[maybe_unused]] int g_n = 0, g_e = 0, g_top = 0, g_kind = 0;
  [[maybe_unused]] GroupLds<GW>* gl = nullptr;
  [[maybe_unused]] const double* g_pt = nullptr;  // this group's pair terms (tensor interaction)
  if constexpr (GRP) {
    // group blockIdx.x on a workgroup of GW waves, element e on lane e of
    // wave 0; lanes of wave 0 past the group's elements leave at once (a
    // barrier waits for waves, not lanes)
    __shared__ GroupLds<GW> lds;
    gl = &lds;
    if (blockIdx.x >= g.group_n) return;
    gp = g.group + blockIdx.x;
    g_n = gp->n_elements;
    if (gp->interaction == CMTJ_INTERACTION_TENSOR && g.pair)
      g_pt = g.pair + (size_t)blockIdx.x * 12 * (size_t)g_n * (size_t)g_n;
    if constexpr (GW > 1) {
      if (threadIdx.x >= 64) {
        group_helper<GW>(lds, g_n, (int)(threadIdx.x >> 6), g_pt);
        return;
      }
    }
    g_e = (int)threadIdx.x;
    if (g_e >= g_n) return;
    g_top = gp->top[g_e];
    g_kind = gp->interaction;
    lane = blockIdx.x * (uint32_t)g_n + (uint32_t)g_e;
  }
  [[maybe_unused]] const cmtj_stack* sk = nullptr;
  [[maybe_unused]] int s_j = 0, s_base = 0;
  if constexpr (STK) {
    // stack s on lanes [s pad, s pad + members) of one wave; the members of
    // all stacks are the group's instances in the caller's order
    const uint32_t pad = g.stack_pad, sid = lane / pad;
    if (sid >= g.stack_n) return;
    sk = g.stack + sid;
    s_j = (int)(lane % pad);
    if (s_j >= sk->n_junctions) return;
    s_base = (int)(threadIdx.x & ~(pad - 1));
    lane = sid * (uint32_t)sk->n_junctions + (uint32_t)s_j;
  }
  if (lane >= g.n) return;
  const uint32_t fl = g.flags;
  const int nl = g.nl;
  const uint32_t jid = g.order[lane];
  const cmtj_junction* J = g.js + jid;
  const uint32_t nsteps = J->n_steps;
  const uint32_t W = J->write_every;
  const uint32_t NE = J->n_entries;
  const double dt = J->dt;
  const uint64_t seed = J->seed, stream = J->stream;
  const uint64_t lbase = J->log_offset;
  const bool logging = (g.log != nullptr) && g.log_mask != 0;
  const uint32_t esize = g.log_f32 ? 4u : 8u;
  char* const lp = logging ? (char*)g.log + lbase * esize : nullptr;
  const uint64_t cs = (uint64_t)NE * esize;
  const double Rp = J->Rp, Rap = J->Rap;
  MRCoef mcoef[LM < kMrRegLayers ? LM : kMrRegLayers];
#pragma unroll
  for (int l = 0; l < (LM < kMrRegLayers ? LM : kMrRegLayers); ++l) mcoef[l] = mr_coef(g, J, l, logging && l < nl);
  const int sps = (SOLVER == CMTJ_RK4) ? 3 : (SOLVER == CMTJ_HEUN ? 2 : 1);

  constexpr bool ANCHORED = (sizeof(R) == 4) && (SOLVER == CMTJ_RK4);
  LP<R> P[LM];
  DV<R> base[LM];
  LP<double> Pd[ANCHORED ? LM : 1];
  DV<double> based[ANCHORED ? LM : 1];
  V3<double> m[LM];
#pragma unroll
  for (int l = 0; l < LM; ++l) {
    if (l < nl) {
      load_layer<R>(g.layers[J->layer_offset + l], g.c, P[l], base[l], m[l]);
      if constexpr (ANCHORED) load_layer<double>(g.layers[J->layer_offset + l], g.c, Pd[l], based[l], m[l]);
    } else {
      dv_zero(base[l]);
      m[l] = mk<double>(0, 0, 0);
      P[l] = P[0];
      if constexpr (ANCHORED) {
        dv_zero(based[l]);
        Pd[l] = Pd[0];
      }
    }
  }
  DynRec dr[DREG];
#pragma unroll
  for (int k = 0; k < DREG; ++k) {
    if (k < g.ndyn)
      dr[k] = load_rec<FM>(g.dyn + (size_t)k * g.n + lane);
    else
      dr[k] = DynRec{0.0, 0.0, 0.0, 0.0, 0, 0};
  }

  FieldWeights<R, LM> fw;
  field_weights<R, LM>(fw, g, 0, dr);
  FieldWeights<float, LM> fwf;
  if constexpr (ANCHORED) field_weights<float, LM>(fwf, g, 0, dr);
  FieldWeights<double, LM> fwd;
  if constexpr (ANCHORED) field_weights<double, LM>(fwd, g, 0, dr);
  ERec<R> er[LM];
  DV<R> dv[LM];
  const R dtR = (R)dt;
  const R sdt = (R)sqrt(dt);
  const R sixth_inv = R(1.0 / 6.0);
  const V3<R> zero = mk<R>(0, 0, 0);
  uint32_t k = 0;
  int code = CMTJ_OK, nan_layer = -1;
  uint32_t i = 0, iend = nsteps;
  cmtj_state* const S = g.state ? g.state + jid : nullptr;
  if (S) {  // resume: state of the previous slice
#pragma unroll
    for (int l = 0; l < LM; ++l)
      if (l < nl) m[l] = mk<double>(S->m[l][0], S->m[l][1], S->m[l][2]);
    i = S->step;
    k = S->entries;
    code = S->code;
    nan_layer = (code == CMTJ_OK) ? -1 : S->nan_layer;
    iend = (code != CMTJ_OK) ? i : (uint32_t)min((uint64_t)nsteps, (uint64_t)i + g.slice);
  }
  uint64_t next_log = (uint64_t)k * W;
  [[maybe_unused]] int s_n = 0, s_ti = 0, s_bi = 0;
  [[maybe_unused]] StackLane sl;
  if constexpr (STK) {
    s_n = sk->n_junctions;
    s_ti = sk->top[s_j];
    s_bi = sk->bottom[s_j];
    sl = stack_lane(sk, s_j, s_n);
  }
  [[maybe_unused]] int g_j1 = 0;
  if constexpr (GRP) {
    int j0_;
    group_share<GW>(g_n, 0, j0_, g_j1);
    const cmtj_layer& Ly = g.layers[J->layer_offset + g_top];
    const double V = Ly.thickness * Ly.cell_surface;
    gl->pos[g_e][0] = gp->pos[g_e][0];
    gl->pos[g_e][1] = gp->pos[g_e][1];
    gl->pos[g_e][2] = gp->pos[g_e][2];
    const int kind = gp->interaction;
    gl->s[g_e] = (kind == CMTJ_INTERACTION_KANAO)    ? (Ly.Ms / g.c.mu0) * V
                 : (kind == CMTJ_INTERACTION_NOMURA) ? Ly.Ms * V / g.c.mu0 / (4 * M_PI)
                                                     : 0.0;  // nullDipoleInteraction
    __syncthreads();
    for (int j = 0; j < g_n; ++j) {
      const double dx = gl->pos[g_e][0] - gl->pos[j][0], dy = gl->pos[g_e][1] - gl->pos[j][1],
                   dz = gl->pos[g_e][2] - gl->pos[j][2];
      // the host has checked |r| against the reference's limits
      gl->il[j][g_e] = (j == g_e) ? 0.0 : 1.0 / sqrt(dx * dx + dy * dy + dz * dz);
    }
  }
  for (; i < iend; ++i) {
    const double t = (double)i * dt;
    [[maybe_unused]] V3<double> gH = mk<double>(0, 0, 0);
    if constexpr (GRP) {
      V3<double> mt = m[0];
#pragma unroll
      for (int l = 0; l < LM; ++l) mt = (l == g_top) ? m[l] : mt;
      __syncthreads();  // the previous step's reads are done
      gl->m[g_e][0] = mt.x;
      gl->m[g_e][1] = mt.y;
      gl->m[g_e][2] = mt.z;
      if (g_e == 0) gl->ctl = 1;
      __syncthreads();  // m and ctl are written
      if (g_kind != CMTJ_INTERACTION_NULL)
        gH = g_pt ? group_field_tensor<GW>(*gl, g_pt, g_n, 0, g_j1, g_e) : group_field<GW>(*gl, 0, g_j1, g_e);
      if constexpr (GW > 1) {
        __syncthreads();  // the other waves' partial fields are written
        if (g_kind != CMTJ_INTERACTION_NULL)
          for (int w = 1; w < GW; ++w)
            gH = gH + mk<double>(gl->part[w - 1][g_e][0], gl->part[w - 1][g_e][1], gl->part[w - 1][g_e][2]);
      }
    }
    [[maybe_unused]] double sI[3] = {0.0, 0.0, 0.0};
    [[maybe_unused]] double sIlog = 0.0, sClog = 0.0;
    // sub-steps of the step: one, or one per member of a non-delayed stack
    // (member q commits at sub-step q; every member computes each sub-step)
    int reps = 1;
    [[maybe_unused]] double s_eff = 1.0;
    if constexpr (STK) {
      if (sl.nd)
        reps = s_n;
      else
        stack_current(sl, s_j, s_n, s_base, 1, stack_d<LM>(sl, s_ti, s_bi, m), t, dt, g.tables, i, sI, sIlog, sClog);
    }
    // one sub-step: the whole step, or member sq's turn in a non-delayed
    // stack (act: this lane commits); true when the run stops here
    auto sub_step = [&](bool act) __attribute__((always_inline)) -> bool {
    if constexpr (ANCHORED) {
      // RK4 of the reference (core/junction.hpp:1017-1041, neighbours frozen
      // at the start of the step) in anchored binary32: k1 and the anchor
      // quantities in binary64, stages 2-4 as binary32 increments against
      // the anchor (rhs_delta), combination m + k1 + (2dk2 + 2dk3 + dk4)/6.
      V3<double> nbBd[LM], nbTd[LM];
      V3<float> nbB[LM], nbT[LM];
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        nbBd[l] = (l > 0) ? m[l > 0 ? l - 1 : 0] : mk<double>(0, 0, 0);
        nbTd[l] = (l + 1 < nl) ? m[l + 1 < LM ? l + 1 : l] : mk<double>(0, 0, 0);
        nbB[l] = cvt<float>(nbBd[l]);
        nbT[l] = cvt<float>(nbTd[l]);
      }
      DV<double> dv0[LM];
      double v0[DREG];
      eval_drivers<double, LM, FM, false, true>(dv0, based, dr, fwd, g, lane, fl, t, i, 0, sps, 0, v0);
      V3<double> k1[LM];
      Anchor A[LM];
      ERec<double> erd;
      DV<float> ds0[LM];
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          k1[l] = rhs<double, FM, false, true>(Pd[l], g.lbits[l], fl, m[l], dv0[l], nbBd[l], nbTd[l], erd, &A[l]) * dt;
          ds0[l] = dv_cvt<float>(dv0[l]);
        }
      }
      DV<float> dd[LM];
      eval_driver_deltas<LM, FM>(dd, dr, fwf, g, lane, fl, t, t + 0.5 * dt, i, 1, sps, v0);
      V3<float> acc[LM], dk[LM];
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          const DV<float> ds = dv_add(ds0[l], dd[l]);
          const V3<float> d2 = cvt<float>(k1[l] * 0.5);
          dk[l] = rhs_delta<FM>(P[l], g.lbits[l], fl, A[l], d2, dd[l], ds, nbB[l], nbT[l]) * dtR;
          acc[l] = dk[l] * 2.f;
          const V3<float> d3 = d2 + dk[l] * 0.5f;
          dk[l] = rhs_delta<FM>(P[l], g.lbits[l], fl, A[l], d3, dd[l], ds, nbB[l], nbT[l]) * dtR;
          acc[l] = acc[l] + dk[l] * 2.f;
        }
      }
      eval_driver_deltas<LM, FM>(dd, dr, fwf, g, lane, fl, t, t + dt, i, 2, sps, v0);
      V3<double> mn[LM];
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          const DV<float> ds = dv_add(ds0[l], dd[l]);
          const V3<float> d4 = cvt<float>(k1[l]) + dk[l];
          const V3<float> dk4 = rhs_delta<FM>(P[l], g.lbits[l], fl, A[l], d4, dd[l], ds, nbB[l], nbT[l]) * dtR;
          acc[l] = acc[l] + dk4;
          mn[l] = normalize_m<R>(m[l] + k1[l] + cvt<double>(acc[l] * sixth_inv));
        }
      }
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl && nan_layer < 0 && act) {
          m[l] = mn[l];
          if (isnan(mn[l].x)) nan_layer = l;
        }
      }
      if (nan_layer >= 0) {
        code = CMTJ_ERR_NAN;
        return true;
      }
    } else if (SOLVER == CMTJ_RK4) {
      // rk4_step (core/junction.hpp:1017-1041) for every layer, neighbours
      // frozen at the start of the step (runMultiLayerSolver, :1682-1707).
      // Stages run outermost so the driver values of one time serve all
      // layers; layers are independent within a step.
      V3<R> nbB[LM], nbT[LM], ms[LM], acc[LM];
#pragma unroll
      for (int l = 0; l < LM; ++l) ms[l] = cvt<R>(m[l]);
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        nbB[l] = (l > 0) ? ms[l > 0 ? l - 1 : 0] : zero;
        nbT[l] = (l + 1 < nl) ? ms[l + 1 < LM ? l + 1 : l] : zero;
      }
      eval_drivers<R, LM, FM, LOGE>(dv, base, dr, fw, g, lane, fl, t, i, 0, sps);
      if constexpr (STK) set_current<R, LM>(dv, sI[0]);
      if constexpr (GRP) add_field_to<R, LM>(dv, g_top, gH);
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          const V3<R> k1 = rhs<R, FM, LOGE>(P[l], g.lbits[l], fl, ms[l], dv[l], nbB[l], nbT[l], er[l]) * dtR;
          acc[l] = k1;
          ms[l] = cvt<R>(m[l]) + k1 * R(0.5);
        }
      }
      eval_drivers<R, LM, FM, LOGE>(dv, base, dr, fw, g, lane, fl, t + 0.5 * dt, i, 1, sps);
      if constexpr (STK) set_current<R, LM>(dv, sI[1]);
      if constexpr (GRP) add_field_to<R, LM>(dv, g_top, gH);
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          const V3<R> k2 = rhs<R, FM, LOGE>(P[l], g.lbits[l], fl, ms[l], dv[l], nbB[l], nbT[l], er[l]) * dtR;
          acc[l] = acc[l] + k2 * R(2.0);
          ms[l] = cvt<R>(m[l]) + k2 * R(0.5);
        }
      }
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          const V3<R> k3 = rhs<R, FM, LOGE>(P[l], g.lbits[l], fl, ms[l], dv[l], nbB[l], nbT[l], er[l]) * dtR;
          acc[l] = acc[l] + k3 * R(2.0);
          ms[l] = cvt<R>(m[l]) + k3;
        }
      }
      eval_drivers<R, LM, FM, LOGE>(dv, base, dr, fw, g, lane, fl, t + dt, i, 2, sps);
      if constexpr (STK) set_current<R, LM>(dv, sI[2]);
      if constexpr (GRP) add_field_to<R, LM>(dv, g_top, gH);
      V3<double> mn[LM];
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          const V3<R> k4 = rhs<R, FM, LOGE>(P[l], g.lbits[l], fl, ms[l], dv[l], nbB[l], nbT[l], er[l]) * dtR;
          acc[l] = acc[l] + k4;
          V3<R> inc;
          if (sizeof(R) == 8)
            inc = mk<R>(acc[l].x / R(6.0), acc[l].y / R(6.0), acc[l].z / R(6.0));
          else
            inc = acc[l] * sixth_inv;
          mn[l] = normalize_m<R>(m[l] + cvt<double>(inc));
        }
      }
      // the reference updates layers in turn; the first NaN stops the run
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl && nan_layer < 0 && act) {
          m[l] = mn[l];
          if (isnan(mn[l].x)) nan_layer = l;
        }
      }
      if constexpr (GRP) {
        // a NaN ends the group's run (the reference throws): the elements
        // are one wave, its lanes leave the loop together
        if (__any(nan_layer >= 0)) {
          if (nan_layer >= 0) code = CMTJ_ERR_NAN;
          return true;
        }
      } else if constexpr (STK) {
        // a member's NaN ends its stack's run (the reference throws): every
        // member of that stack stops at this step (a non-delayed stack's later
        // members keep their start-of-step m, as in the reference), the lanes
        // of the stack leave the loop together, other stacks go on
        const unsigned long long bad = __ballot(nan_layer >= 0);
        const unsigned long long mine =
            ((g.stack_pad >= 64) ? ~0ull : ((1ull << g.stack_pad) - 1)) << (unsigned)s_base;
        if (bad & mine) {
          if (nan_layer >= 0) code = CMTJ_ERR_NAN;
          return true;
        }
      } else if (nan_layer >= 0) {
        code = CMTJ_ERR_NAN;
        return true;
      }
    } else if (SOLVER == CMTJ_EULER_HEUN) {
      // eulerHeunSolverStep, core/junction.hpp:1709-1754
      V3<R> mr_[LM];
#pragma unroll
      for (int l = 0; l < LM; ++l) mr_[l] = cvt<R>(m[l]);
      eval_drivers<R, LM, FM, LOGE>(dv, base, dr, fw, g, lane, fl, t, i, 0, sps);
      if constexpr (STK) set_current<R, LM>(dv, sI[0]);
      if constexpr (GRP) add_field_to<R, LM>(dv, g_top, gH);
      V3<double> mn[LM];
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          const V3<R> bot = (l > 0) ? mr_[l > 0 ? l - 1 : 0] : zero;
          const V3<R> top = (l + 1 < nl) ? mr_[l + 1 < LM ? l + 1 : l] : zero;
          const V3<R> f = rhs<R, FM, LOGE>(P[l], g.lbits[l], fl, mr_[l], dv[l], bot, top, er[l]);
          V3<double> acc = m[l] + cvt<double>(f * dtR);
          if (CMTJ_HAS(F_THERMAL) && ((g.lbits[l] >> 6) & 1)) {
            const R sigma = sqrt(P[l].sig2 * dv[l].T);
            const V3<R> dW = thermal_normals<R>(seed, stream, i, (uint32_t)l) * sigma;
            const V3<R> gn = stoch<R>(P[l], mr_[l], dW);
            const V3<R> mN = mr_[l] + gn * sdt;
            const V3<R> gp = stoch<R>(P[l], mN, dW);
            acc = acc + cvt<double>((gn + gp) * R(0.5) * sdt);
          }
          mn[l] = acc;
        }
      }
#pragma unroll
      for (int l = 0; l < LM; ++l)
        if (l < nl && act) m[l] = normalize_m<R>(mn[l]);
    } else {
      // heunSolverStep, core/junction.hpp:1756-1815
      V3<R> mr_[LM], fn[LM], gn[LM], dW[LM], mN[LM];
#pragma unroll
      for (int l = 0; l < LM; ++l) mr_[l] = cvt<R>(m[l]);
      eval_drivers<R, LM, FM, LOGE>(dv, base, dr, fw, g, lane, fl, t, i, 0, sps);
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          const V3<R> bot = (l > 0) ? mr_[l > 0 ? l - 1 : 0] : zero;
          const V3<R> top = (l + 1 < nl) ? mr_[l + 1 < LM ? l + 1 : l] : zero;
          fn[l] = rhs<R, FM, LOGE>(P[l], g.lbits[l], fl, mr_[l], dv[l], bot, top, er[l]);
          dW[l] = zero;
          gn[l] = zero;
          if (CMTJ_HAS(F_THERMAL) && ((g.lbits[l] >> 6) & 1)) {
            const R sigma = sqrt(P[l].sig2 * dv[l].T);
            dW[l] = thermal_normals<R>(seed, stream, i, (uint32_t)l) * sigma;
            gn[l] = stoch<R>(P[l], mr_[l], dW[l]);
          }
          mN[l] = mr_[l] + fn[l] * dtR + gn[l] * sdt;
        }
      }
      eval_drivers<R, LM, FM, LOGE>(dv, base, dr, fw, g, lane, fl, t + dt, i, 1, sps);
#pragma unroll
      for (int l = 0; l < LM; ++l) {
        if (l < nl) {
          const V3<R> bot = (l > 0) ? mN[l > 0 ? l - 1 : 0] : zero;
          const V3<R> top = (l + 1 < nl) ? mN[l + 1 < LM ? l + 1 : l] : zero;
          const V3<R> f2 = rhs<R, FM, LOGE>(P[l], g.lbits[l], fl, mN[l], dv[l], bot, top, er[l]);
          V3<double> acc = m[l] + cvt<double>((fn[l] + f2) * (R(0.5) * dtR));
          if (CMTJ_HAS(F_THERMAL) && ((g.lbits[l] >> 6) & 1)) {
            const V3<R> g2 = stoch<R>(P[l], mN[l], dW[l]);
            acc = acc + cvt<double>((gn[l] + g2) * R(0.5) * sdt);
          }
          if (act) m[l] = normalize_m<R>(acc);
        }
      }
    }
      return false;
    };
    if constexpr (STK) {
      bool halt = false;
      for (int sq = 0; sq < reps && !halt; ++sq) {
        if (sl.nd) stack_current_nd<LM>(sl, s_j, s_n, s_base, s_ti, s_bi, m, t, dt, sq, s_eff, sI, sIlog, sClog);
        halt = sub_step(!sl.nd || s_j == sq);
      }
      if (halt) break;
    } else {
      if (sub_step(true)) break;
    }
    if ((uint64_t)i == next_log) {
      if (logging) write_entry<R, LM, FM, LOGE>(g, J, m, er, lp + (uint64_t)k * esize, cs, Rp, Rap, mcoef);
      if constexpr (STK) {
        if (logging) {  // the member's stack channels follow its junction channels
          const uint64_t nch = (uint64_t)log_channels(nl, g.mr_mode, g.log_mask);
          char* q = lp + (uint64_t)k * esize + nch * cs;
          log_put(g.log_f32, q, sIlog);
          log_put(g.log_f32, q + cs, sClog);
        }
      }
      ++k;
      next_log += W;
    }
  }
  if constexpr (GRP && GW > 1) {
    // release the other waves: one more round of the first two barriers
    // with ctl = 0
    __syncthreads();
    if (g_e == 0) gl->ctl = 0;
    __syncthreads();
  }
  if (g.mag_out) {
#pragma unroll
    for (int l = 0; l < LM; ++l) {
      if (l < nl) {
        double* o = g.mag_out + ((size_t)jid * CMTJ_MAX_LAYERS + l) * 3;
        o[0] = m[l].x;
        o[1] = m[l].y;
        o[2] = m[l].z;
      }
    }
  }
  if (g.status) {
    cmtj_status s;
    s.code = code;
    s.steps_done = i;
    s.entries_done = k;
    s.nan_layer = nan_layer;
    g.status[jid] = s;
  }
  if (S) {
#pragma unroll
    for (int l = 0; l < LM; ++l) {
      if (l < nl) {
        S->m[l][0] = m[l].x;
        S->m[l][1] = m[l].y;
        S->m[l][2] = m[l].z;
      }
    }
    S->step = i;
    S->entries = k;
    S->code = code;
    S->nan_layer = nan_layer;
  }
}

// One instance per lane.  (A variant compiled for >= 4 waves per SIMD spilled
// 88-300 VGPRs and ran 1.6-5x slower on full-chip batches; not kept.)
template <class R, int LM, int SOLVER, uint32_t FM, bool LOGE, bool STK = false, bool GRP = false>
__global__ __launch_bounds__(64) void k_junction(GroupArgs g) {
  junction_body<R, LM, SOLVER, FM, LOGE, STK, GRP>(g);
}

// Dipole-coupled groups on GW waves each: wave 0 integrates the elements,
// waves 1..GW-1 share the sum over the sources.
template <class R, int LM, int SOLVER, uint32_t FM, int GW>
__global__ __launch_bounds__(64 * GW) void k_group(GroupArgs g) {
  junction_body<R, LM, SOLVER, FM, false, false, true, GW>(g);
}

}  // namespace cmtj
