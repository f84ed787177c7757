// kernels_dp.h -- Dormand-Prince 5(4) with adaptive steps, one instance per
// lane (FP64): k_dp for one layer, k_dp_ml for 2..LM layers.
//
// Each lane runs its own step-size history: Layer::dormand_prince_step
// (core/junction.hpp:866-1002, the standard controller) inside the adaptive
// branch of Junction::runSimulation (:1995-2044).  Quirks of the reference
// kept on purpose, since they define its results:
//   * the clock advances by the step size proposed for the *next* step
//     (dormand_prince_step overwrites timeStep before `t += current_timestep`);
//   * the rejection counter is reset at the top of every attempt, so a step
//     is forced through only when the proposed step falls below 1e-15 s;
//   * a log entry follows each accepted step that reaches the next multiple
//     of the write frequency, labelled with the new time.
// Lanes of a wave take different numbers of attempts; every lane leaves
// its loop after at most n_steps attempts (a run that needs more ends with
// CMTJ_ERR_LIMIT), so the wave always drains.
#pragma once

#include "kernels.h"

namespace cmtj {

__device__ __forceinline__ double vlen(V3<double> v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

// Step-size control of one layer: AdaptiveIntegrationParams
// (core/junction.hpp:125-139) as dormand_prince_step applies it (:942-982).
// The PID controller's integral and previous error live in the layer and
// advance on every attempt, accepted or rejected.
struct DpCtl {
  double atol, rtol, fmaxv, fminv, safety, kp, ki, kd, integral, prev;
  bool pid;
  __device__ void load(const cmtj_layer& Ly) {
    atol = Ly.dp_abs_tol;
    rtol = Ly.dp_rel_tol;
    fmaxv = Ly.dp_max_factor;
    fminv = Ly.dp_min_factor;
    safety = Ly.dp_safety;
    pid = Ly.dp_use_pid != 0;
    kp = Ly.dp_kp;
    ki = Ly.dp_ki;
    kd = Ly.dp_kd;
    integral = Ly.dp_integral_error;
    prev = Ly.dp_prev_error_ratio;
  }
  __device__ void unit() {
    atol = rtol = fmaxv = fminv = safety = 1.0;
    kp = ki = kd = integral = prev = 0.0;
    pid = false;
  }
  // error ratio of an attempt from |m5 - m4| and |m_t| (:942-950)
  __device__ double ratio(double err, double mlen) const {
    const double tol = atol + rtol * mlen;
    double r = err / tol;
    if (tol < 1e-15 || isnan(r) || isinf(r)) r = 10.0;
    return r;
  }
  // the step proposed after an attempt of step h with error ratio r (:952-982)
  __device__ double propose(double h, double r) {
    double factor;
    if (pid) {
      // the exponent's roundings spelled out (two fma, one product), so
      // every kernel that inlines this rounds it alike
      const double e = log(r);
      integral += e;
      factor = exp(fma(kp, e, fma(ki, integral, kd * (e - prev))));
      prev = e;
    } else {
      factor = safety * pow(1.0 / r, 0.2);
    }
    if (isnan(factor) || isinf(factor)) factor = fminv;
    factor = fmin(factor, fmaxv);
    factor = fmax(factor, fminv);
    double nh = h * factor;
    if (isnan(nh) || isinf(nh) || nh <= 0) nh = h * fminv;
    return nh;
  }
};

// The final controller state of the layers after their magnetisations in
// an instance's mag_out block (only when a layer runs the PID controller).
template <int LM>
__device__ __forceinline__ void store_pid(double* o, int nl, const DpCtl* c) {
  bool any = false;
#pragma unroll
  for (int l = 0; l < LM; ++l) any |= (l < nl) && c[l].pid;
  if (!any) return;
#pragma unroll
  for (int l = 0; l < LM; ++l)
    if (l < nl) {
      o[3 * nl + 2 * l] = c[l].integral;
      o[3 * nl + 2 * l + 1] = c[l].prev;
    }
}

template <uint32_t FM>
__global__ __launch_bounds__(64) void k_dp(GroupArgs g) {
  const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= g.n) return;
  const uint32_t fl = g.flags;
  const uint32_t jid = g.order[lane];
  const cmtj_junction* J = g.js + jid;
  const cmtj_layer& Ly = g.layers[J->layer_offset];
  LP<double> P;
  DV<double> base[1];
  V3<double> m[1];
  load_layer<double>(Ly, g.c, P, base[0], m[0]);
  DynRec dr[DREG];
#pragma unroll
  for (int k = 0; k < DREG; ++k) {
    if (k < g.ndyn)
      dr[k] = load_rec<FM>(g.dyn + (size_t)k * g.n + lane);
    else
      dr[k] = DynRec{0.0, 0.0, 0.0, 0.0, 0, 0};
  }
  FieldWeights<double, 1> fw;
  field_weights<double, 1>(fw, g, 0, dr);
  ERec<double> er[1];
  const int bits = g.lbits[0];
  const V3<double> zero = mk<double>(0, 0, 0);
  // dm/dt at (t, m) (calculateLLGWithFieldTorque, :851-864)
  auto f = [&](double t, V3<double> mm) -> V3<double> {
    DV<double> dv[1];
    eval_drivers<double, 1, FM, false>(dv, base, dr, fw, g, lane, fl, t, 0, 0, 1);
    return rhs<double, FM, false>(P, bits, fl, mm, dv[0], zero, zero, er[0]);
  };
  const double T = J->total_time, wf = J->write_freq, h0 = J->dt;
  DpCtl ctl[1];
  ctl[0].load(Ly);
  const uint32_t cap = J->n_steps, NE = J->n_entries;
  const bool logging = (g.log != nullptr) && g.log_mask != 0;
  const MRCoef mcoef[1] = {mr_coef(g, J, 0, logging)};
  const uint32_t esize = g.log_f32 ? 4u : 8u;
  char* const lp = logging ? (char*)g.log + J->log_offset * esize : nullptr;
  const uint64_t cs = (uint64_t)NE * esize;
  const uint64_t tch = (uint64_t)(log_channels(1, g.mr_mode, g.log_mask) - 1);  // the time channel is last

  double t = 0.0, h = h0, next_write = 0.0;
  uint32_t it = 0, k = 0, accepted_steps = 0;
  int code = CMTJ_OK, nan_layer = -1;
  while (t < T) {
    if (it >= cap) {
      code = CMTJ_ERR_LIMIT;
      break;
    }
    ++it;
    if (t + h > T) h = T - t;
    int rejections = 0;
    if (isnan(h) || isinf(h) || h <= 0) {
      h = h0;
      ++rejections;
    }
    // dormand_prince_step(t, h, 0, 0)
    const V3<double> mt = m[0];
    const V3<double> k1 = f(t, mt) * h;
    const V3<double> k2 = f(t + (1.0 / 5.0) * h, mt + k1 * (1.0 / 5.0)) * h;
    const V3<double> k3 = f(t + (3.0 / 10.0) * h, mt + k1 * (3.0 / 40.0) + k2 * (9.0 / 40.0)) * h;
    const V3<double> k4 =
        f(t + (4.0 / 5.0) * h, mt + k1 * (44.0 / 45.0) + k2 * (-56.0 / 15.0) + k3 * (32.0 / 9.0)) * h;
    const V3<double> k5 = f(t + (8.0 / 9.0) * h, mt + k1 * (19372.0 / 6561.0) + k2 * (-25360.0 / 2187.0) +
                                                      k3 * (64448.0 / 6561.0) + k4 * (-212.0 / 729.0)) *
                          h;
    const V3<double> k6 = f(t + 1.0 * h, mt + k1 * (9017.0 / 3168.0) + k2 * (-355.0 / 33.0) +
                                             k3 * (46732.0 / 5247.0) + k4 * (49.0 / 176.0) +
                                             k5 * (-5103.0 / 18656.0)) *
                          h;
    const V3<double> k7 = f(t + 1.0 * h, mt + k1 * (35.0 / 384.0) + k3 * (500.0 / 1113.0) +
                                             k4 * (125.0 / 192.0) + k5 * (-2187.0 / 6784.0) + k6 * (11.0 / 84.0)) *
                          h;
    V3<double> m5 = mt + k1 * (35.0 / 384.0) + k3 * (500.0 / 1113.0) + k4 * (125.0 / 192.0) +
                    k5 * (-2187.0 / 6784.0) + k6 * (11.0 / 84.0);
    const V3<double> m4 = mt + k1 * (5179.0 / 57600.0) + k3 * (7571.0 / 16695.0) + k4 * (393.0 / 640.0) +
                          k5 * (-92097.0 / 339200.0) + k6 * (187.0 / 2100.0) + k7 * (1.0 / 40.0);
    const double ratio = ctl[0].ratio(vlen(m5 - m4), vlen(mt));
    const double nh = ctl[0].propose(h, ratio);
    bool acc;
    if (ratio <= 1.0) {
      m5 = normalize_exact(m5);
      if (isnan(m5.x)) {  // the reference throws "NAN magnetisation"
        m[0] = m5;
        code = CMTJ_ERR_NAN;
        nan_layer = 0;
        break;
      }
      m[0] = m5;
      h = nh;
      acc = true;
    } else {
      h = nh;
      acc = false;
    }
    if (!acc && (h < 1e-15 || rejections >= 10)) acc = true;  // forced acceptance (:2017-2025)
    if (acc) {
      t += h;
      ++accepted_steps;
      if (t >= next_write) {
        if (k >= NE) {
          code = CMTJ_ERR_LIMIT;
          break;
        }
        if (logging) {
          char* p = lp + (uint64_t)k * esize;
          write_entry<double, 1, FM, false>(g, J, m, er, p, cs, J->Rp, J->Rap, mcoef);
          log_put(g.log_f32, p + tch * cs, t);
        }
        ++k;
        next_write = (double)((uint32_t)(t / wf) + 1u) * wf;
      }
    }
  }
  if (g.mag_out) {
    double* o = g.mag_out + (size_t)jid * CMTJ_MAX_LAYERS * 3;
    o[0] = m[0].x;
    o[1] = m[0].y;
    o[2] = m[0].z;
    store_pid<1>(o, 1, ctl);
  }
  if (g.status) {
    cmtj_status s;
    s.code = code;
    s.steps_done = accepted_steps;
    s.entries_done = k;
    s.nan_layer = nan_layer;
    g.status[jid] = s;
  }
}

// Several layers (Junction::runMultiLayerSolver, core/junction.hpp:1682-1707
// inside the adaptive loop :1995-2044).  Each attempt freezes the
// neighbours at the start-of-attempt magnetisations; the layers step in
// order, each with the step size the previous layer proposed (every
// dormand_prince_step writes its proposal into the one shared step), and a
// layer that accepts keeps its new m whatever the others do.  The clock
// advances by the last layer's proposal.  A rejection anywhere makes the
// reference set the step to the smallest Layer::hopt, which stays at its
// initial -1 (:259): the attempt is forced through with a step of -1 s and
// the clock runs backwards, after which that run does not finish.  The lane
// stops there with CMTJ_ERR_DP_REJECTED, keeping the attempt's
// magnetisations.  Layer parameters stay in registers (the per-layer loops
// unroll over LM); driver values are evaluated for every layer at each
// stage time (the layers' stage times differ).
template <int LM, uint32_t FM>
__global__ __launch_bounds__(64) void k_dp_ml(GroupArgs g) {
  const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane >= g.n) return;
  const uint32_t fl = g.flags;
  const int nl = g.nl;
  const uint32_t jid = g.order[lane];
  const cmtj_junction* J = g.js + jid;
  LP<double> P[LM];
  DV<double> base[LM];
  V3<double> m[LM];
  DpCtl ctl[LM];
#pragma unroll
  for (int l = 0; l < LM; ++l) {
    if (l < nl) {
      const cmtj_layer& Ly = g.layers[J->layer_offset + l];
      load_layer<double>(Ly, g.c, P[l], base[l], m[l]);
      ctl[l].load(Ly);
    } else {
      dv_zero(base[l]);
      m[l] = mk<double>(0, 0, 0);
      P[l] = P[0];
      ctl[l].unit();
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
  FieldWeights<double, LM> fw;
  field_weights<double, LM>(fw, g, 0, dr);
  ERec<double> er[LM];
  const V3<double> zero = mk<double>(0, 0, 0);
  const double T = J->total_time, wf = J->write_freq, h0 = J->dt;
  const uint32_t cap = J->n_steps, NE = J->n_entries;
  const bool logging = (g.log != nullptr) && g.log_mask != 0;
  MRCoef mcoef[LM < kMrRegLayers ? LM : kMrRegLayers];
#pragma unroll
  for (int l = 0; l < (LM < kMrRegLayers ? LM : kMrRegLayers); ++l) mcoef[l] = mr_coef(g, J, l, logging && l < nl);
  const uint32_t esize = g.log_f32 ? 4u : 8u;
  char* const lp = logging ? (char*)g.log + J->log_offset * esize : nullptr;
  const uint64_t cs = (uint64_t)NE * esize;
  const uint64_t tch = (uint64_t)(log_channels(nl, g.mr_mode, g.log_mask) - 1);  // the time channel is last

  double t = 0.0, h = h0, next_write = 0.0;
  uint32_t it = 0, k = 0, accepted_steps = 0;
  int code = CMTJ_OK, nan_layer = -1;
  while (t < T) {
    if (it >= cap) {
      code = CMTJ_ERR_LIMIT;
      break;
    }
    ++it;
    if (t + h > T) h = T - t;
    if (isnan(h) || isinf(h) || h <= 0) h = h0;
    V3<double> cp[LM];  // start-of-attempt magnetisations (magCopies)
#pragma unroll
    for (int l = 0; l < LM; ++l) cp[l] = m[l];
    bool acc = true;
#pragma unroll
    for (int l = 0; l < LM; ++l) {
      if (l < nl && code == CMTJ_OK) {
        const V3<double> bot = (l > 0) ? cp[l > 0 ? l - 1 : 0] : zero;
        const V3<double> top = (l + 1 < nl) ? cp[l + 1 < LM ? l + 1 : l] : zero;
        // dm/dt of layer l at (ts, mm) (calculateLLGWithFieldTorque, :851-864)
        auto f = [&](double ts, V3<double> mm) -> V3<double> {
          DV<double> dv[LM];
          eval_drivers<double, LM, FM, false>(dv, base, dr, fw, g, lane, fl, ts, 0, 0, 1);
          return rhs<double, FM, false>(P[l], g.lbits[l], fl, mm, dv[l], bot, top, er[l]);
        };
        const V3<double> mt = m[l];
        const V3<double> k1 = f(t, mt) * h;
        const V3<double> k2 = f(t + (1.0 / 5.0) * h, mt + k1 * (1.0 / 5.0)) * h;
        const V3<double> k3 = f(t + (3.0 / 10.0) * h, mt + k1 * (3.0 / 40.0) + k2 * (9.0 / 40.0)) * h;
        const V3<double> k4 =
            f(t + (4.0 / 5.0) * h, mt + k1 * (44.0 / 45.0) + k2 * (-56.0 / 15.0) + k3 * (32.0 / 9.0)) * h;
        const V3<double> k5 = f(t + (8.0 / 9.0) * h, mt + k1 * (19372.0 / 6561.0) + k2 * (-25360.0 / 2187.0) +
                                                          k3 * (64448.0 / 6561.0) + k4 * (-212.0 / 729.0)) *
                              h;
        const V3<double> k6 = f(t + 1.0 * h, mt + k1 * (9017.0 / 3168.0) + k2 * (-355.0 / 33.0) +
                                                 k3 * (46732.0 / 5247.0) + k4 * (49.0 / 176.0) +
                                                 k5 * (-5103.0 / 18656.0)) *
                              h;
        const V3<double> k7 = f(t + 1.0 * h, mt + k1 * (35.0 / 384.0) + k3 * (500.0 / 1113.0) +
                                                 k4 * (125.0 / 192.0) + k5 * (-2187.0 / 6784.0) + k6 * (11.0 / 84.0)) *
                              h;
        V3<double> m5 = mt + k1 * (35.0 / 384.0) + k3 * (500.0 / 1113.0) + k4 * (125.0 / 192.0) +
                        k5 * (-2187.0 / 6784.0) + k6 * (11.0 / 84.0);
        const V3<double> m4 = mt + k1 * (5179.0 / 57600.0) + k3 * (7571.0 / 16695.0) + k4 * (393.0 / 640.0) +
                              k5 * (-92097.0 / 339200.0) + k6 * (187.0 / 2100.0) + k7 * (1.0 / 40.0);
        const double ratio = ctl[l].ratio(vlen(m5 - m4), vlen(mt));
        h = ctl[l].propose(h, ratio);  // the next layer steps with this layer's proposal
        if (ratio <= 1.0) {
          m5 = normalize_exact(m5);
          m[l] = m5;
          if (isnan(m5.x)) {  // the reference throws "NAN magnetisation"
            code = CMTJ_ERR_NAN;
            nan_layer = l;
          }
        } else {
          acc = false;
        }
      }
    }
    if (code != CMTJ_OK) break;
    if (!acc) {
      code = CMTJ_ERR_DP_REJECTED;
      break;
    }
    t += h;
    ++accepted_steps;
    if (t >= next_write) {
      if (k >= NE) {
        code = CMTJ_ERR_LIMIT;
        break;
      }
      if (logging) {
        char* p = lp + (uint64_t)k * esize;
        write_entry<double, LM, FM, false>(g, J, m, er, p, cs, J->Rp, J->Rap, mcoef);
        log_put(g.log_f32, p + tch * cs, t);
      }
      ++k;
      next_write = (double)((uint32_t)(t / wf) + 1u) * wf;
    }
  }
  if (g.mag_out) {
    double* o = g.mag_out + (size_t)jid * CMTJ_MAX_LAYERS * 3;
#pragma unroll
    for (int l = 0; l < LM; ++l)
      if (l < nl) {
        o[3 * l + 0] = m[l].x;
        o[3 * l + 1] = m[l].y;
        o[3 * l + 2] = m[l].z;
      }
    store_pid<LM>(o, nl, ctl);
  }
  if (g.status) {
    cmtj_status s;
    s.code = code;
    s.steps_done = accepted_steps;
    s.entries_done = k;
    s.nan_layer = nan_layer;
    g.status[jid] = s;
  }
}

// Single-layer Dormand-Prince from a device work queue: a persistent grid
// smaller than the batch, in which each lane runs an instance to its end and
// then claims the next position of the group's order (`queue`, zeroed before
// the launch) until the order is exhausted.  Adaptive runs take step counts
// no one knows in advance, so lanes of a wave finish at different times; the
// loop is flat (one attempt per pass, a lane whose instance ended claims and
// loads the next one in the same pass), so a lane never waits for the other
// lanes' instances to end.  Per instance the arithmetic is k_dp's, so the
// results are k_dp's bit for bit.
template <uint32_t FM>
__global__ __launch_bounds__(64) void k_dp_q(GroupArgs g, uint32_t* queue) {
  const uint32_t fl = g.flags;
  const int bits = g.lbits[0];
  const bool logging = (g.log != nullptr) && g.log_mask != 0;
  const uint32_t esize = g.log_f32 ? 4u : 8u;
  const uint64_t tch = (uint64_t)(log_channels(1, g.mr_mode, g.log_mask) - 1);  // the time channel is last
  const V3<double> zero = mk<double>(0, 0, 0);
  // the current instance
  uint32_t pos = 0, jid = 0;
  const cmtj_junction* J = nullptr;
  LP<double> P;
  DV<double> base[1];
  V3<double> m[1];
  DynRec dr[DREG];
  FieldWeights<double, 1> fw;
  ERec<double> er[1];
  MRCoef mcoef[1];
  double T = 0, wf = 1, h0 = 0;
  DpCtl ctl[1];
  ctl[0].unit();
  uint32_t cap = 0, NE = 0;
  char* lp = nullptr;
  uint64_t cs = 0;
  double t = 0.0, h = 0.0, next_write = 0.0;
  uint32_t it = 0, k = 0, accepted_steps = 0;
  int code = CMTJ_OK, nan_layer = -1;

  auto load = [&](uint32_t q) {
    pos = q;
    jid = g.order[q];
    J = g.js + jid;
    const cmtj_layer& Ly = g.layers[J->layer_offset];
    load_layer<double>(Ly, g.c, P, base[0], m[0]);
#pragma unroll
    for (int r = 0; r < DREG; ++r) {
      if (r < g.ndyn)
        dr[r] = load_rec<FM>(g.dyn + (size_t)r * g.n + q);
      else
        dr[r] = DynRec{0.0, 0.0, 0.0, 0.0, 0, 0};
    }
    field_weights<double, 1>(fw, g, 0, dr);
    T = J->total_time;
    wf = J->write_freq;
    h0 = J->dt;
    ctl[0].load(Ly);
    cap = J->n_steps;
    NE = J->n_entries;
    mcoef[0] = mr_coef(g, J, 0, logging);
    lp = logging ? (char*)g.log + J->log_offset * esize : nullptr;
    cs = (uint64_t)NE * esize;
    t = 0.0;
    h = h0;
    next_write = 0.0;
    it = k = accepted_steps = 0;
    code = CMTJ_OK;
    nan_layer = -1;
  };
  auto store = [&]() {
    if (g.mag_out) {
      double* o = g.mag_out + (size_t)jid * CMTJ_MAX_LAYERS * 3;
      o[0] = m[0].x;
      o[1] = m[0].y;
      o[2] = m[0].z;
      store_pid<1>(o, 1, ctl);
    }
    if (g.status) {
      cmtj_status s;
      s.code = code;
      s.steps_done = accepted_steps;
      s.entries_done = k;
      s.nan_layer = nan_layer;
      g.status[jid] = s;
    }
  };
  // dm/dt at (ts, mm) (calculateLLGWithFieldTorque, :851-864)
  auto f = [&](double ts, V3<double> mm) -> V3<double> {
    DV<double> dv[1];
    eval_drivers<double, 1, FM, false>(dv, base, dr, fw, g, pos, fl, ts, 0, 0, 1);
    return rhs<double, FM, false>(P, bits, fl, mm, dv[0], zero, zero, er[0]);
  };
  // one pass of the adaptive loop (:1995-2044); true when the run ended
  auto attempt = [&]() -> bool {
    if (!(t < T)) return true;
    if (it >= cap) {
      code = CMTJ_ERR_LIMIT;
      return true;
    }
    ++it;
    if (t + h > T) h = T - t;
    int rejections = 0;
    if (isnan(h) || isinf(h) || h <= 0) {
      h = h0;
      ++rejections;
    }
    const V3<double> mt = m[0];
    const V3<double> k1 = f(t, mt) * h;
    const V3<double> k2 = f(t + (1.0 / 5.0) * h, mt + k1 * (1.0 / 5.0)) * h;
    const V3<double> k3 = f(t + (3.0 / 10.0) * h, mt + k1 * (3.0 / 40.0) + k2 * (9.0 / 40.0)) * h;
    const V3<double> k4 =
        f(t + (4.0 / 5.0) * h, mt + k1 * (44.0 / 45.0) + k2 * (-56.0 / 15.0) + k3 * (32.0 / 9.0)) * h;
    const V3<double> k5 = f(t + (8.0 / 9.0) * h, mt + k1 * (19372.0 / 6561.0) + k2 * (-25360.0 / 2187.0) +
                                                      k3 * (64448.0 / 6561.0) + k4 * (-212.0 / 729.0)) *
                          h;
    const V3<double> k6 = f(t + 1.0 * h, mt + k1 * (9017.0 / 3168.0) + k2 * (-355.0 / 33.0) +
                                             k3 * (46732.0 / 5247.0) + k4 * (49.0 / 176.0) +
                                             k5 * (-5103.0 / 18656.0)) *
                          h;
    const V3<double> k7 = f(t + 1.0 * h, mt + k1 * (35.0 / 384.0) + k3 * (500.0 / 1113.0) +
                                             k4 * (125.0 / 192.0) + k5 * (-2187.0 / 6784.0) + k6 * (11.0 / 84.0)) *
                          h;
    V3<double> m5 = mt + k1 * (35.0 / 384.0) + k3 * (500.0 / 1113.0) + k4 * (125.0 / 192.0) +
                    k5 * (-2187.0 / 6784.0) + k6 * (11.0 / 84.0);
    const V3<double> m4 = mt + k1 * (5179.0 / 57600.0) + k3 * (7571.0 / 16695.0) + k4 * (393.0 / 640.0) +
                          k5 * (-92097.0 / 339200.0) + k6 * (187.0 / 2100.0) + k7 * (1.0 / 40.0);
    const double ratio = ctl[0].ratio(vlen(m5 - m4), vlen(mt));
    const double nh = ctl[0].propose(h, ratio);
    bool acc;
    if (ratio <= 1.0) {
      m5 = normalize_exact(m5);
      m[0] = m5;
      if (isnan(m5.x)) {  // the reference throws "NAN magnetisation"
        code = CMTJ_ERR_NAN;
        nan_layer = 0;
        return true;
      }
      h = nh;
      acc = true;
    } else {
      h = nh;
      acc = false;
    }
    if (!acc && (h < 1e-15 || rejections >= 10)) acc = true;  // forced acceptance (:2017-2025)
    if (acc) {
      t += h;
      ++accepted_steps;
      if (t >= next_write) {
        if (k >= NE) {
          code = CMTJ_ERR_LIMIT;
          return true;
        }
        if (logging) {
          char* p = lp + (uint64_t)k * esize;
          write_entry<double, 1, FM, false>(g, J, m, er, p, cs, J->Rp, J->Rap, mcoef);
          log_put(g.log_f32, p + tch * cs, t);
        }
        ++k;
        next_write = (double)((uint32_t)(t / wf) + 1u) * wf;
      }
    }
    return false;
  };

  uint32_t q = atomicAdd(queue, 1u);
  bool have = q < g.n;
  if (have) load(q);
  while (__ballot(have)) {
    if (have && attempt()) {
      store();
      q = atomicAdd(queue, 1u);
      have = q < g.n;
      if (have) load(q);
    }
  }
}

// Clears the work counters of a launch (a kernel, not a memset: the plan's
// launches may be captured into a graph).
__global__ void k_clear_u32(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace cmtj
