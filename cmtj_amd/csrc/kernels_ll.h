// kernels_ll.h -- the junction kernel with one layer per lane.
//
// An instance of nl layers occupies L consecutive lanes (L = 2, 4 or 8, a
// power of two >= nl), lane l integrating layer l.  The layers of a
// multilayer step are independent once the neighbours are fixed (RK4 and
// Euler-Heun use the start-of-step neighbours, core/junction.hpp:1682-1707
// and 1709-1754; Heun's corrector uses the predicted neighbours,
// :1756-1815), so a step needs one exchange of m between adjacent lanes
// (two for Heun).  The exchange is a DPP row shift: L divides 16, so an
// instance never straddles a DPP row.  Per wave this divides the
// field/torque work of a step by nl, which is what bounds a batch too small
// to fill the chip: the time of a sweep is the step count of its longest
// instance times the instruction latency of one step of one wave.
//
// Arithmetic per layer is the one of k_junction (same functions, same
// order), so the results are bit-identical to it.
#pragma once

#include "kernels.h"

namespace cmtj {

// v of lane (i - 1) [CTRL = row_shr:1] or lane (i + 1) [row_shl:1]; lanes
// whose source is outside the 16-lane row read 0.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  int lo = (int)(unsigned)(u & 0xffffffffull);
  int hi = (int)(unsigned)(u >> 32);
  lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, true);
  hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
constexpr int DPP_ROW_SHR1 = 0x111;
constexpr int DPP_ROW_SHL1 = 0x101;

template <class R>
__device__ __forceinline__ V3<R> from_prev(V3<R> v, bool on) {
  const V3<double> d = cvt<double>(v);
  const V3<double> r = mk<double>(dpp_d<DPP_ROW_SHR1>(d.x), dpp_d<DPP_ROW_SHR1>(d.y), dpp_d<DPP_ROW_SHR1>(d.z));
  return on ? cvt<R>(r) : mk<R>(0, 0, 0);
}
template <class R>
__device__ __forceinline__ V3<R> from_next(V3<R> v, bool on) {
  const V3<double> d = cvt<double>(v);
  const V3<double> r = mk<double>(dpp_d<DPP_ROW_SHL1>(d.x), dpp_d<DPP_ROW_SHL1>(d.y), dpp_d<DPP_ROW_SHL1>(d.z));
  return on ? cvt<R>(r) : mk<R>(0, 0, 0);
}

// Sum of v over the L lanes of an instance (butterfly; for two layers the
// result equals the sequential sum exactly).
template <int L>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = 1; o < L; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// logLayerParams (core/junction.hpp:1580-1636) for entry k, written by the
// lanes of one instance: each lane its own layer's m (and energy) channels,
// lane 0 the resistance channels.  top = m of the next layer after the step.
template <class R, int L, uint32_t FM, bool LOGE>
__device__ __forceinline__ void write_entry_ll(const GroupArgs& g, const cmtj_junction* J, int ly, bool act,
                                               V3<double> m, V3<double> top, const ERec<R>& er, char* p,
                                               uint64_t cs, double Rp, double Rap, const MRCoef& mc) {
  const int nl = g.nl, lmask = g.log_mask, mr = g.mr_mode;
  const bool f32 = g.log_f32;
  const int nm = ((lmask & CMTJ_LOG_MX) ? 1 : 0) + ((lmask & CMTJ_LOG_MYZ) ? 2 : 0);
  if (act && nm) {
    char* q = p + (uint64_t)(ly * nm) * cs;
    if (lmask & CMTJ_LOG_MX) {
      log_put(f32, q, m.x);
      q += cs;
    }
    if (lmask & CMTJ_LOG_MYZ) {
      log_put(f32, q, m.y);
      q += cs;
      log_put(f32, q, m.z);
    }
  }
  p += (uint64_t)(nl * nm) * cs;
  if (lmask & CMTJ_LOG_MR) {
    if (mr == CMTJ_MR_CLASSIC) {
      // calculateMagnetoresistance (:1865-1867): layer 0 against layer 1
      // (one layer: against the fixed reference vector, :1623-1629)
      if (ly == 0) {
        double cosT;
        if (nl == 1) {
          const double* rf = mc.ref;
          cosT = m.x * rf[0] + m.y * rf[1] + m.z * rf[2];
        } else {
          cosT = m.x * top.x + m.y * top.y + m.z * top.z;
        }
        log_put(f32, p, Rp + (((Rap - Rp) / 2.0) * (1.0 - cosT)));
      }
      p += cs;
    } else if (mr == CMTJ_MR_STRIP) {
      // stripMagnetoResistance (:1835-1857), parallel sum of the layers
      double rx = 0.0, ry = 0.0;
      if (act) {
        const double Rx = mc.rx0 + mc.amrx * (m.x * m.x) + mc.smrx * (m.y * m.y);
        const double Ry = mc.ry0 + mc.ahe2 * m.z + mc.xy * m.x * m.y;
        rx = 1. / Rx;
        ry = 1. / Ry;
      }
      rx = group_sum<L>(rx);
      ry = group_sum<L>(ry);
      if (ly == 0) {
        log_put(f32, p, 1 / rx);
        log_put(f32, p + cs, 1 / ry);
      }
      p += 2 * cs;
    }
  }
  if (LOGE && (lmask & CMTJ_LOG_ENERGY) && act) {
    char* q = p + (uint64_t)(ly * CMTJ_ENERGY_CHANNELS) * cs;
    const ERec<R>& e = er;
    const double v[CMTJ_ENERGY_CHANNELS] = {
        e.K, e.K2, e.Jb, e.Jt, e.I, e.Hext.x, e.Hext.y, e.Hext.z,
        e.Hiec.x, e.Hiec.y, e.Hiec.z, e.Hanis.x, e.Hanis.y, e.Hanis.z,
        e.Hdemag.x, e.Hdemag.y, e.Hdemag.z, e.Hfl.x, e.Hfl.y, e.Hfl.z,
        e.Hdl.x, e.Hdl.y, e.Hdl.z};
#pragma unroll
    for (int c = 0; c < CMTJ_ENERGY_CHANNELS; ++c) {
      log_put(f32, q, v[c]);
      q += cs;
    }
  }
}

// WPE: waves per SIMD the register allocation must leave room for (1: no
// limit).  A batch that fills the chip many times over hides latency with
// more waves; 4 caps the kernel at 128 VGPRs (the CIMS ensemble's kernel
// spills 216 B to scratch at that cap and still runs 1.22x faster).
template <class R, int L, int SOLVER, uint32_t FM, bool LOGE, int WPE = 1>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_junction_ll(GroupArgs g) {
  static_assert(L == 2 || L == 4 || L == 8, "L lanes per instance must divide a 16-lane DPP row");
  const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t inst = gl / L;
  if (inst >= g.n) return;
  const int ly = (int)(gl % L);
  const uint32_t fl = g.flags;
  const int nl = g.nl;
  const bool act = ly < nl;
  const int lyr = act ? ly : 0;  // idle lanes shadow layer 0 and write nothing
  const uint32_t jid = g.order[inst];
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
  const MRCoef mcoef = mr_coef(g, J, lyr, logging && act);
  const int sps = (SOLVER == CMTJ_RK4) ? 3 : (SOLVER == CMTJ_HEUN ? 2 : 1);
  const bool hasB = act && ly > 0, hasT = act && ly + 1 < nl;
  const uint32_t gshift = (threadIdx.x & 63u) & ~(uint32_t)(L - 1);  // first lane of the instance in the wave

  int bits = 0;
#pragma unroll
  for (int l = 0; l < L; ++l) bits = (lyr == l) ? g.lbits[l] : bits;

  constexpr bool ANCHORED = (sizeof(R) == 4) && (SOLVER == CMTJ_RK4);
  LP<R> P;
  DV<R> base[1];
  LP<double> Pd;
  DV<double> based[1];
  V3<double> m;
  load_layer<R>(g.layers[J->layer_offset + lyr], g.c, P, base[0], m);
  if constexpr (ANCHORED) load_layer<double>(g.layers[J->layer_offset + lyr], g.c, Pd, based[0], m);
  DynRec dr[DREG];
#pragma unroll
  for (int k = 0; k < DREG; ++k) {
    if (k < g.ndyn)
      dr[k] = load_rec<FM>(g.dyn + (size_t)k * g.n + inst);
    else
      dr[k] = DynRec{0.0, 0.0, 0.0, 0.0, 0, 0};
  }
  FieldWeights<R, 1> fw;
  field_weights<R, 1>(fw, g, lyr, dr);
  FieldWeights<float, 1> fwf;
  if constexpr (ANCHORED) field_weights<float, 1>(fwf, g, lyr, dr);
  FieldWeights<double, 1> fwd;
  if constexpr (ANCHORED) field_weights<double, 1>(fwd, g, lyr, dr);
  ERec<R> er;
  DV<R> dv[1];
  const R dtR = (R)dt;
  const R sdt = (R)sqrt(dt);
  const R sixth_inv = R(1.0 / 6.0);
  const V3<R> zero = mk<R>(0, 0, 0);
  uint32_t k = 0;
  int code = CMTJ_OK, nan_layer = -1;
  uint32_t i = 0, iend = nsteps;
  cmtj_state* const S = g.state ? g.state + jid : nullptr;
  if (S) {  // resume: state of the previous slice
    m = mk<double>(S->m[lyr][0], S->m[lyr][1], S->m[lyr][2]);
    i = S->step;
    k = S->entries;
    code = S->code;
    nan_layer = (code == CMTJ_OK) ? -1 : S->nan_layer;
    iend = (code != CMTJ_OK) ? i : (uint32_t)min((uint64_t)nsteps, (uint64_t)i + g.slice);
  }
  uint64_t next_log = (uint64_t)k * W;
  // neighbours (m of layers ly - 1 and ly + 1), refreshed after every step
  V3<double> bot = from_prev(m, hasB), top = from_next(m, hasT);
  for (; i < iend; ++i) {
    const double t = (double)i * dt;
    V3<double> mn;
    if constexpr (ANCHORED) {
      // anchored binary32 RK4, as in k_junction
      const V3<float> nbB = cvt<float>(bot), nbT = cvt<float>(top);
      DV<double> dv0[1];
      double v0[DREG];
      eval_drivers<double, 1, FM, false, true>(dv0, based, dr, fwd, g, inst, fl, t, i, 0, sps, lyr, v0);
      Anchor A;
      ERec<double> erd;
      const V3<double> k1 = rhs<double, FM, false, true>(Pd, bits, fl, m, dv0[0], bot, top, erd, &A) * dt;
      const DV<float> ds0 = dv_cvt<float>(dv0[0]);
      DV<float> dd[1];
      eval_driver_deltas<1, FM>(dd, dr, fwf, g, inst, fl, t, t + 0.5 * dt, i, 1, sps, v0, lyr);
      V3<float> acc, dk;
      {
        const DV<float> ds = dv_add(ds0, dd[0]);
        const V3<float> d2 = cvt<float>(k1 * 0.5);
        dk = rhs_delta<FM>(P, bits, fl, A, d2, dd[0], ds, nbB, nbT) * dtR;
        acc = dk * 2.f;
        const V3<float> d3 = d2 + dk * 0.5f;
        dk = rhs_delta<FM>(P, bits, fl, A, d3, dd[0], ds, nbB, nbT) * dtR;
        acc = acc + dk * 2.f;
      }
      eval_driver_deltas<1, FM>(dd, dr, fwf, g, inst, fl, t, t + dt, i, 2, sps, v0, lyr);
      {
        const DV<float> ds = dv_add(ds0, dd[0]);
        const V3<float> d4 = cvt<float>(k1) + dk;
        const V3<float> dk4 = rhs_delta<FM>(P, bits, fl, A, d4, dd[0], ds, nbB, nbT) * dtR;
        acc = acc + dk4;
        mn = normalize_m<R>(m + k1 + cvt<double>(acc * sixth_inv));
      }
    } else if (SOLVER == CMTJ_RK4) {
      // rk4_step (core/junction.hpp:1017-1041), neighbours of the start of
      // the step (runMultiLayerSolver, :1682-1707)
      const V3<R> nbB = cvt<R>(bot), nbT = cvt<R>(top);
      V3<R> ms = cvt<R>(m), acc;
      eval_drivers<R, 1, FM, LOGE>(dv, base, dr, fw, g, inst, fl, t, i, 0, sps, lyr);
      const V3<R> k1 = rhs<R, FM, LOGE>(P, bits, fl, ms, dv[0], nbB, nbT, er) * dtR;
      acc = k1;
      ms = cvt<R>(m) + k1 * R(0.5);
      eval_drivers<R, 1, FM, LOGE>(dv, base, dr, fw, g, inst, fl, t + 0.5 * dt, i, 1, sps, lyr);
      const V3<R> k2 = rhs<R, FM, LOGE>(P, bits, fl, ms, dv[0], nbB, nbT, er) * dtR;
      acc = acc + k2 * R(2.0);
      ms = cvt<R>(m) + k2 * R(0.5);
      const V3<R> k3 = rhs<R, FM, LOGE>(P, bits, fl, ms, dv[0], nbB, nbT, er) * dtR;
      acc = acc + k3 * R(2.0);
      ms = cvt<R>(m) + k3;
      eval_drivers<R, 1, FM, LOGE>(dv, base, dr, fw, g, inst, fl, t + dt, i, 2, sps, lyr);
      const V3<R> k4 = rhs<R, FM, LOGE>(P, bits, fl, ms, dv[0], nbB, nbT, er) * dtR;
      acc = acc + k4;
      V3<R> inc;
      if (sizeof(R) == 8)
        inc = mk<R>(acc.x / R(6.0), acc.y / R(6.0), acc.z / R(6.0));
      else
        inc = acc * sixth_inv;
      mn = normalize_m<R>(m + cvt<double>(inc));
    } else if (SOLVER == CMTJ_EULER_HEUN) {
      // eulerHeunSolverStep, core/junction.hpp:1709-1754
      const V3<R> mr_ = cvt<R>(m);
      eval_drivers<R, 1, FM, LOGE>(dv, base, dr, fw, g, inst, fl, t, i, 0, sps, lyr);
      const V3<R> f = rhs<R, FM, LOGE>(P, bits, fl, mr_, dv[0], cvt<R>(bot), cvt<R>(top), er);
      V3<double> acc = m + cvt<double>(f * dtR);
      if (CMTJ_HAS(F_THERMAL) && ((bits >> 6) & 1)) {
        const R sigma = sqrt(P.sig2 * dv[0].T);
        const V3<R> dW = thermal_normals<R>(seed, stream, i, (uint32_t)lyr) * sigma;
        const V3<R> gn = stoch<R>(P, mr_, dW);
        const V3<R> mN = mr_ + gn * sdt;
        const V3<R> gp = stoch<R>(P, mN, dW);
        acc = acc + cvt<double>((gn + gp) * R(0.5) * sdt);
      }
      mn = normalize_m<R>(acc);
    } else {
      // heunSolverStep, core/junction.hpp:1756-1815; the corrector sees the
      // predicted neighbours
      const V3<R> mr_ = cvt<R>(m);
      eval_drivers<R, 1, FM, LOGE>(dv, base, dr, fw, g, inst, fl, t, i, 0, sps, lyr);
      const V3<R> fn = rhs<R, FM, LOGE>(P, bits, fl, mr_, dv[0], cvt<R>(bot), cvt<R>(top), er);
      V3<R> dW = zero, gn = zero;
      if (CMTJ_HAS(F_THERMAL) && ((bits >> 6) & 1)) {
        const R sigma = sqrt(P.sig2 * dv[0].T);
        dW = thermal_normals<R>(seed, stream, i, (uint32_t)lyr) * sigma;
        gn = stoch<R>(P, mr_, dW);
      }
      const V3<R> mN = mr_ + fn * dtR + gn * sdt;
      const V3<R> bN = from_prev(mN, hasB), tN = from_next(mN, hasT);
      eval_drivers<R, 1, FM, LOGE>(dv, base, dr, fw, g, inst, fl, t + dt, i, 1, sps, lyr);
      const V3<R> f2 = rhs<R, FM, LOGE>(P, bits, fl, mN, dv[0], bN, tN, er);
      V3<double> acc = m + cvt<double>((fn + f2) * (R(0.5) * dtR));
      if (CMTJ_HAS(F_THERMAL) && ((bits >> 6) & 1)) {
        const V3<R> g2 = stoch<R>(P, mN, dW);
        acc = acc + cvt<double>((gn + g2) * R(0.5) * sdt);
      }
      mn = normalize_m<R>(acc);
    }
    if (SOLVER == CMTJ_RK4) {
      // the reference updates layers in turn and stops at the first NaN:
      // layers up to and including it take the new value
      const unsigned long long bad = __ballot(act && isnan(mn.x));
      const uint32_t gb = (uint32_t)((bad >> gshift) & ((1ull << L) - 1));
      if (gb) {
        nan_layer = __builtin_ctz(gb);
        if (ly <= nan_layer) m = mn;
        code = CMTJ_ERR_NAN;
        break;
      }
    }
    m = mn;
    bot = from_prev(m, hasB);
    top = from_next(m, hasT);
    if ((uint64_t)i == next_log) {
      if (logging)
        write_entry_ll<R, L, FM, LOGE>(g, J, ly, act, m, top, er, lp + (uint64_t)k * esize, cs, Rp, Rap, mcoef);
      ++k;
      next_log += W;
    }
  }
  if (g.mag_out && act) {
    double* o = g.mag_out + ((size_t)jid * CMTJ_MAX_LAYERS + ly) * 3;
    o[0] = m.x;
    o[1] = m.y;
    o[2] = m.z;
  }
  if (g.status && ly == 0) {
    cmtj_status s;
    s.code = code;
    s.steps_done = i;
    s.entries_done = k;
    s.nan_layer = nan_layer;
    g.status[jid] = s;
  }
  if (S) {
    if (act) {
      S->m[ly][0] = m.x;
      S->m[ly][1] = m.y;
      S->m[ly][2] = m.z;
    }
    if (ly == 0) {
      S->step = i;
      S->entries = k;
      S->code = code;
      S->nan_layer = nan_layer;
    }
  }
}

}  // namespace cmtj
