/*
 * oracle.c -- CPU FP64 restatement of cmtj's (s)LLGS time integration.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never part of the product path).
 *
 * Plain C, scalar binary64, compiled without FP contraction so every product
 * and sum rounds separately, as the reference's x86-64 build does.  Each
 * function restates one reference function; the reference line is cited.
 * Thermal noise: the reference draws std::normal_distribution over an
 * unseedable mt19937 (core/junction.hpp:174-175); this oracle and the engine
 * share a counter-based Philox4x32-10 + Box-Muller stream instead, so a
 * stochastic run is reproducible and comparable trajectory by trajectory.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double x, y, z;
} vec3;

static vec3 v3(double x, double y, double z) {
  vec3 r = {x, y, z};
  return r;
}
static vec3 vadd(vec3 a, vec3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static vec3 vsub(vec3 a, vec3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static vec3 vmul(vec3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
static vec3 smul(double s, vec3 a) { return v3(s * a.x, s * a.y, s * a.z); }
static vec3 vdiv(vec3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
/* c_dot, core/junction.hpp:82-84 */
static double vdot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* c_cross, core/junction.hpp:74-80 */
static vec3 vcross(vec3 a, vec3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static vec3 vload(const double* p) { return v3(p[0], p[1], p[2]); }
/* CVector::length / normalize, core/cvector.hpp:179-192 */
static double vlength(vec3 a) { return sqrt(pow(a.x, 2) + pow(a.y, 2) + pow(a.z, 2)); }
static vec3 vnormalize(vec3 a) {
  const double mag = vlength(a);
  if (mag != 0) {
    a.x = a.x / mag;
    a.y = a.y / mag;
    a.z = a.z / mag;
  }
  return a;
}

/* calculate_tensor_interaction, core/junction.hpp:54-62 */
static vec3 tensor_interaction(vec3 m, const double* t, double Ms, double mu0) {
  vec3 r = v3(t[0] * m.x + t[1] * m.y + t[2] * m.z, t[3] * m.x + t[4] * m.y + t[5] * m.z,
              t[6] * m.x + t[7] * m.y + t[8] * m.z);
  return vmul(r, Ms / mu0);
}

/* ---- drivers ------------------------------------------------------------ */

/* ScalarDriver::getCurrentScalarValue, core/drivers.hpp:262-305, with the
 * helpers stepUpdate (:73-78), pulseTrain (:79-88), trapezoidalUpdate
 * (:90-108).  CMTJ_TABLE returns the pre-sampled callback value. */
double oracle_driver_value(const cmtj_driver* d, double time, uint32_t step, int stage,
                           int samples_per_step, const double* tables) {
  double v = d->constant;
  switch (d->type) {
    case CMTJ_PULSE: {
      const int n = (int)(time / d->period);
      const double dT = d->cycle * d->period;
      const double nT = n * d->period;
      if (nT <= time && time <= (nT + dT)) v += d->amplitude;
      break;
    }
    case CMTJ_SINE:
      v += d->amplitude * sin(2 * M_PI * time * d->frequency + d->phase);
      break;
    case CMTJ_POSINE:
      v += fabs(d->amplitude * sin(2 * M_PI * time * d->frequency + d->phase));
      break;
    case CMTJ_HALFSINE: {
      const double tamp = d->amplitude * sin(2 * M_PI * time * d->frequency + d->phase);
      if (tamp <= 0) v += tamp;
      break;
    }
    case CMTJ_STEP:
      if (time >= d->t_start && time <= d->t_stop) v += d->amplitude;
      break;
    case CMTJ_TRAPEZOID: {
      const double a = d->amplitude, t0 = d->t_start, e = d->edge_time, s = d->steady_time;
      double add = 0;
      if (time < t0)
        add = 0;
      else if (time <= t0 + e)
        add = (a / e) * (time - t0);
      else if (time <= t0 + e + s)
        add = a;
      else if (time <= t0 + 2 * e + s)
        add = a - (a / e) * (time - (t0 + e + s));
      v += add;
      break;
    }
    case CMTJ_GAUSSIMPULSE:
      v += d->amplitude * exp(-pow(time - d->t_start, 2) / (2 * pow(d->edge_time, 2)));
      break;
    case CMTJ_GAUSSSTEP:
      v += 0.5 * d->amplitude * (1 + erf((time - d->t_start) / (sqrt(2) * d->edge_time)));
      break;
    case CMTJ_TABLE:
      /* the reference returns the callback value alone (drivers.hpp:294-302) */
      return tables[(size_t)d->table_index + (size_t)step * (size_t)samples_per_step + (size_t)stage];
    default:
      break;
  }
  return v;
}

/* ---- Philox4x32-10 (Salmon et al., SC'11) + Box-Muller ------------------ */

static void philox_round(uint32_t c[4], const uint32_t k[2]) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0];
  const uint32_t n1 = lo1;
  const uint32_t n2 = hi0 ^ c[3] ^ k[1];
  const uint32_t n3 = lo0;
  c[0] = n0;
  c[1] = n1;
  c[2] = n2;
  c[3] = n3;
}

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r) {
    philox_round(c, k);
    if (r < 9) {
      k[0] += 0x9E3779B9u;
      k[1] += 0xBB67AE85u;
    }
  }
  memcpy(out, c, sizeof(c));
}

/* Counter {step, layer, stream lo, stream hi}, key {seed lo, seed hi};
 * u = (x + 0.5) 2^-32; xi = sqrt(-2 ln u0) (cos 2pi u1, sin 2pi u1),
 * sqrt(-2 ln u2) cos 2pi u3. */
void oracle_thermal_normals(uint64_t seed, uint64_t stream, uint32_t step, uint32_t layer,
                            double xi[3]) {
  const uint32_t ctr[4] = {step, layer, (uint32_t)stream, (uint32_t)(stream >> 32)};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  oracle_philox4x32_10(ctr, key, o);
  const double s = 2.3283064365386963e-10; /* 2^-32 */
  const double u0 = ((double)o[0] + 0.5) * s, u1 = ((double)o[1] + 0.5) * s;
  const double u2 = ((double)o[2] + 0.5) * s, u3 = ((double)o[3] + 0.5) * s;
  const double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
  xi[0] = r0 * cos(2.0 * M_PI * u1);
  xi[1] = r0 * sin(2.0 * M_PI * u1);
  xi[2] = r1 * cos(2.0 * M_PI * u3);
}

/* ---- one layer ------------------------------------------------------------ */

typedef struct {
  const cmtj_layer* p;
  const cmtj_constants* c;
  const double* tables;
  int sps; /* table samples per step */
  vec3 mag;
  double cellVolume;
  /* PID controller state of Dormand-Prince (AdaptiveIntegrationParams
   * integral_error, prev_error_ratio; kept by the layer across runs) */
  double pid_integral, pid_prev;
  /* values of the last field/torque evaluation (logged with energies) */
  double K_log, K2_log, Jbottom_log, Jtop_log, I_log;
  vec3 Hext, HIEC, HAnis, Hdemag, Hfl_v, Hdl_v;
} olayer;

static double drv(const olayer* L, int slot, double t, uint32_t step, int stage) {
  return oracle_driver_value(&L->p->drv[slot], t, step, stage, L->sps, L->tables);
}
/* AxialDriver::getCurrentAxialDrivers, core/drivers.hpp:429-433 */
static vec3 axial(const olayer* L, int slot, double t, uint32_t step, int stage) {
  return v3(drv(L, slot, t, step, stage), drv(L, slot + 1, t, step, stage),
            drv(L, slot + 2, t, step, stage));
}

/* calculateIEC_, core/junction.hpp:654-662 */
static vec3 iec_(const olayer* L, double J, double J2, vec3 m, vec3 coupled) {
  return vdiv(vmul(coupled, J + 2 * J2 * vdot(coupled, m)), L->p->Ms * L->p->thickness);
}
/* calculateIDMI_, core/junction.hpp:677-693 */
static vec3 idmi_(const olayer* L, vec3 D, vec3 coupled) {
  return vdiv(smul(-1.0, vcross(D, coupled)), L->p->Ms * L->p->thickness);
}

/* calculateHeff + calculateHeffDipoleInjection, core/junction.hpp:579-622,
 * with the term functions :624-701.  Stores the logged values. */
static vec3 heff(olayer* L, double t, uint32_t step, int stage, vec3 m, vec3 bottom, vec3 top) {
  const cmtj_layer* p = L->p;
  const double mu0 = L->c->mu0;
  const vec3 Hdipole = vadd(tensor_interaction(bottom, p->dipole_bottom, p->Ms, mu0),
                            tensor_interaction(top, p->dipole_top, p->Ms, mu0));
  L->Hext = axial(L, CMTJ_SLOT_HEXT, t, step, stage);
  const vec3 Hoe = axial(L, CMTJ_SLOT_HOE, t, step, stage);
  L->Hdemag = tensor_interaction(m, p->demag, p->Ms, mu0);
  /* calculateIEC, :664-675 */
  L->Jbottom_log = drv(L, CMTJ_SLOT_J_BOTTOM, t, step, stage);
  L->Jtop_log = drv(L, CMTJ_SLOT_J_TOP, t, step, stage);
  const double J2b = drv(L, CMTJ_SLOT_J2_BOTTOM, t, step, stage);
  const double J2t = drv(L, CMTJ_SLOT_J2_TOP, t, step, stage);
  L->HIEC = vadd(iec_(L, L->Jbottom_log, J2b, m, bottom), iec_(L, L->Jtop_log, J2t, m, top));
  /* calculateIDMI, :695-701 */
  const vec3 Hidmi = vadd(idmi_(L, axial(L, CMTJ_SLOT_IDMI_BOTTOM, t, step, stage), bottom),
                          idmi_(L, axial(L, CMTJ_SLOT_IDMI_TOP, t, step, stage), top));
  /* calculateAnisotropy, :638-643 */
  L->K_log = drv(L, CMTJ_SLOT_K, t, step, stage);
  const vec3 anis = vload(p->anis);
  L->HAnis = vmul(anis, (2 * L->K_log) * vdot(anis, m) / (p->Ms));
  const vec3 Hdmi = axial(L, CMTJ_SLOT_HDMI, t, step, stage);
  const vec3 Hres = axial(L, CMTJ_SLOT_HRES, t, step, stage);
  /* calculateSecondOrderAnisotropy, :645-652 */
  L->K2_log = drv(L, CMTJ_SLOT_K2, t, step, stage);
  const vec3 HAnis2 = vmul(anis, (4 * L->K2_log) * pow(vdot(anis, m), 3) / (p->Ms));
  const vec3 Hfluct = v3(0, 0, 0);
  /* summation order of :607-620 */
  vec3 H = vadd(L->Hext, L->HAnis);
  H = vadd(H, HAnis2);
  H = vadd(H, L->HIEC);
  H = vadd(H, Hidmi);
  H = vadd(H, Hoe);
  H = vadd(H, Hdmi);
  H = vadd(H, Hfluct);
  H = vsub(H, L->Hdemag);
  H = vsub(H, Hdipole);
  H = vadd(H, Hres);
  return H;
}

/* solveLLG, core/junction.hpp:716-815 */
static vec3 solve_llg(olayer* L, double t, uint32_t step, int stage, vec3 m, vec3 bottom, vec3 top,
                      vec3 H) {
  const cmtj_layer* p = L->p;
  const double GYRO = L->c->gyro;
  const vec3 prod = vcross(m, H);
  const vec3 prod2 = vcross(m, prod);
  const double convTerm = 1 / (1 + pow(p->damping, 2));
  const vec3 dmdt = vadd(prod, vmul(prod2, p->damping));
  vec3 reference = v3(0, 0, 0);
  switch (p->ref_type) {
    case CMTJ_REF_FIXED:
      reference = vload(p->ref);
      break;
    case CMTJ_REF_TOP:
      reference = top;
      break;
    case CMTJ_REF_BOTTOM:
      reference = bottom;
      break;
    default:
      break;
  }
  if (p->include_stt) {
    L->I_log = drv(L, CMTJ_SLOT_CURRENT, t, step, stage);
    const double aJ = L->c->hbar * L->I_log / (L->c->echarge * p->Ms * p->thickness);
    double eta;
    if (p->alternative_stt) {
      eta = p->polarisation / (1 + p->slonczewski * vdot(m, reference));
    } else {
      const double slonSq = pow(p->slonczewski, 2);
      eta = (p->polarisation * slonSq) / (slonSq + 1 + (slonSq - 1) * vdot(m, reference));
    }
    const double sttTerm = GYRO * aJ * eta;
    const vec3 fieldLike = vcross(m, reference);
    const vec3 dampingLike = vcross(m, fieldLike);
    vec3 r = vadd(vmul(dmdt, -GYRO), vmul(vmul(dampingLike, -sttTerm), p->kappa));
    r = vadd(r, vmul(vmul(fieldLike, sttTerm), p->beta));
    return vmul(r, convTerm);
  } else if (p->include_sot) {
    double Hdl = 0, Hfl = 0, Hdl2 = 0, Hfl2 = 0;
    if (p->dynamic_sot) {
      Hdl = drv(L, CMTJ_SLOT_DL, t, step, stage);
      Hfl = drv(L, CMTJ_SLOT_FL, t, step, stage);
      Hdl2 = drv(L, CMTJ_SLOT_DL2, t, step, stage);
      Hfl2 = drv(L, CMTJ_SLOT_FL2, t, step, stage);
    } else {
      L->I_log = drv(L, CMTJ_SLOT_CURRENT, t, step, stage);
      Hdl = p->dl_torque * L->I_log;
      Hfl = p->fl_torque * L->I_log;
    }
    const vec3 ref2 = vload(p->ref2);
    L->Hfl_v = vmul(reference, Hfl - p->damping * Hdl);
    L->Hdl_v = vmul(reference, Hdl + p->damping * Hfl);
    const vec3 cm = vcross(m, reference);
    const vec3 ccm = vcross(m, cm);
    const vec3 cm2 = vcross(m, ref2);
    const vec3 ccm2 = vcross(m, cm2);
    const vec3 flT = vmul(cm, Hfl - p->damping * Hdl);
    const vec3 dlT = vmul(ccm, Hdl + p->damping * Hfl);
    const vec3 flT2 = vmul(cm2, Hfl2 - p->damping * Hdl2);
    const vec3 dlT2 = vmul(ccm2, Hdl2 + p->damping * Hfl2);
    vec3 r = vadd(vadd(vadd(vadd(dmdt, flT), dlT), flT2), dlT2);
    return vmul(vmul(r, -GYRO), convTerm);
  }
  return vmul(vmul(dmdt, -GYRO), convTerm);
}

/* calculateLLGWithFieldTorque, core/junction.hpp:855-864 */
static vec3 llg(olayer* L, double t, uint32_t step, int stage, vec3 m, vec3 bottom, vec3 top) {
  const vec3 H = heff(L, t, step, stage, m, bottom, top);
  return solve_llg(L, t, step, stage, m, bottom, top, H);
}

/* stochasticTorque, core/junction.hpp:824-831 */
static vec3 stochastic_torque(const olayer* L, vec3 m, vec3 dW) {
  const double convTerm = -L->c->gyro / (1. + pow(L->p->damping, 2));
  const vec3 thcross = vcross(m, dW);
  const vec3 thcross2 = vcross(m, thcross);
  return vmul(vadd(thcross, vmul(thcross2, L->p->damping)), convTerm);
}

/* getLangevinStochasticStandardDeviation + getStochasticLangevinVector,
 * core/junction.hpp:1111-1128 (the draw uses the shared Philox stream). */
static vec3 langevin_vector(const olayer* L, const cmtj_junction* j, double t, uint32_t step,
                            uint32_t layer_index) {
  if (!L->p->temperature_set) return v3(0, 0, 0);
  const double T = drv(L, CMTJ_SLOT_TEMPERATURE, t, step, 0);
  const double mainFactor =
      (2 * L->p->damping * L->c->kB * T) / (L->p->Ms * L->cellVolume * L->c->gyro);
  const double sigma = sqrt(mainFactor);
  double xi[3];
  oracle_thermal_normals(j->seed, j->stream, step, layer_index, xi);
  return vmul(v3(xi[0], xi[1], xi[2]), sigma);
}

/* rk4_step, core/junction.hpp:1017-1041.  Returns 1 on NaN. */
static int rk4_step(olayer* L, double time, double dt, uint32_t step, vec3 bottom, vec3 top) {
  vec3 m_t = L->mag;
  const vec3 k1 = vmul(llg(L, time, step, 0, m_t, bottom, top), dt);
  const vec3 k2 = vmul(llg(L, time + 0.5 * dt, step, 1, vadd(m_t, vmul(k1, 0.5)), bottom, top), dt);
  const vec3 k3 = vmul(llg(L, time + 0.5 * dt, step, 1, vadd(m_t, vmul(k2, 0.5)), bottom, top), dt);
  const vec3 k4 = vmul(llg(L, time + dt, step, 2, vadd(m_t, k3), bottom, top), dt);
  m_t = vadd(m_t, vdiv(vadd(vadd(vadd(k1, vmul(k2, 2.0)), vmul(k3, 2.0)), k4), 6.0));
  m_t = vnormalize(m_t);
  L->mag = m_t;
  return isnan(L->mag.x) ? 1 : 0;
}

int oracle_n_channels(int n_layers, int mr_mode, int log_mask) {
  int n = 0;
  if (log_mask & CMTJ_LOG_MX) n += n_layers;
  if (log_mask & CMTJ_LOG_MYZ) n += 2 * n_layers;
  if (log_mask & CMTJ_LOG_MR) n += (mr_mode == CMTJ_MR_CLASSIC) ? 1 : (mr_mode == CMTJ_MR_STRIP ? 2 : 0);
  if (log_mask & CMTJ_LOG_ENERGY) n += CMTJ_ENERGY_CHANNELS * n_layers;
  if (log_mask & CMTJ_LOG_TIME) n += 1;
  return n;
}

/* logLayerParams (core/junction.hpp:1580-1636) into channel rows, with
 * calculateMagnetoresistance (:1865-1867) and stripMagnetoResistance
 * (:1835-1857). */
static void log_entry(const cmtj_junction* j, const olayer* Ls, double* log, uint32_t k) {
  const int nl = j->n_layers;
  const size_t E = j->n_entries;
  int ch = 0;
  for (int l = 0; l < nl; ++l) {
    if (j->log_mask & CMTJ_LOG_MX) log[(size_t)(ch++) * E + k] = Ls[l].mag.x;
    if (j->log_mask & CMTJ_LOG_MYZ) {
      log[(size_t)(ch++) * E + k] = Ls[l].mag.y;
      log[(size_t)(ch++) * E + k] = Ls[l].mag.z;
    }
  }
  if (j->log_mask & CMTJ_LOG_MR) {
    if (j->mr_mode == CMTJ_MR_CLASSIC) {
      const double cosTheta = (nl == 1) ? vdot(Ls[0].mag, vload(Ls[0].p->ref))
                                        : vdot(Ls[0].mag, Ls[1].mag);
      log[(size_t)(ch++) * E + k] = j->Rp + (((j->Rap - j->Rp) / 2.0) * (1.0 - cosTheta));
    } else if (j->mr_mode == CMTJ_MR_STRIP) {
      double Rx_acc = 0.0, Ry_acc = 0.0;
      for (int l = 0; l < nl; ++l) {
        const vec3 m = Ls[l].mag;
        const double Rx = j->Rx0[l] + j->AMR_X[l] * (m.x * m.x) + j->SMR_X[l] * (m.y * m.y);
        const double Ry = j->Ry0[l] + 0.5 * j->AHE[l] * m.z + (j->AMR_Y[l] + j->SMR_Y[l]) * m.x * m.y;
        Rx_acc += 1. / Rx;
        Ry_acc += 1. / Ry;
      }
      log[(size_t)(ch++) * E + k] = 1 / Rx_acc;
      log[(size_t)(ch++) * E + k] = 1 / Ry_acc;
    }
  }
  if (j->log_mask & CMTJ_LOG_ENERGY) {
    for (int l = 0; l < nl; ++l) {
      const olayer* L = &Ls[l];
      const int sot = L->p->include_sot;
      const double vals[CMTJ_ENERGY_CHANNELS] = {
          L->K_log,  L->K2_log,  L->Jbottom_log, L->Jtop_log, L->I_log,
          L->Hext.x, L->Hext.y,  L->Hext.z,      L->HIEC.x,   L->HIEC.y,
          L->HIEC.z, L->HAnis.x, L->HAnis.y,     L->HAnis.z,  L->Hdemag.x,
          L->Hdemag.y, L->Hdemag.z,
          sot ? L->Hfl_v.x : 0, sot ? L->Hfl_v.y : 0, sot ? L->Hfl_v.z : 0,
          sot ? L->Hdl_v.x : 0, sot ? L->Hdl_v.y : 0, sot ? L->Hdl_v.z : 0,
      };
      for (int q = 0; q < CMTJ_ENERGY_CHANNELS; ++q) log[(size_t)(ch++) * E + k] = vals[q];
    }
  }
}

/* Layer states of one junction at the start of a run. */
static void load_layers(const cmtj_junction* j, const cmtj_layer* layers, const cmtj_constants* c,
                        const double* tables, olayer* Ls) {
  const int nl = j->n_layers;
  memset(Ls, 0, sizeof(olayer) * CMTJ_MAX_LAYERS);
  const int sps = (j->solver == CMTJ_RK4) ? 3 : (j->solver == CMTJ_HEUN ? 2 : 1);
  for (int l = 0; l < nl; ++l) {
    Ls[l].p = &layers[l];
    Ls[l].c = c;
    Ls[l].tables = tables;
    Ls[l].sps = sps;
    Ls[l].mag = vload(layers[l].mag);
    Ls[l].cellVolume = layers[l].cell_surface * layers[l].thickness;
    Ls[l].pid_integral = layers[l].dp_integral_error;
    Ls[l].pid_prev = layers[l].dp_prev_error_ratio;
  }
}

/* One step i of every layer of a junction with its solver.  Returns 0, or
 * nonzero with st->code / st->nan_layer set. */
static int junction_step(const cmtj_junction* j, olayer* Ls, uint32_t i, cmtj_status* st) {
  const int nl = j->n_layers;
  const double dt = j->dt;
  const vec3 zero = v3(0, 0, 0);
  const double t = i * dt; /* core/junction.hpp:1974 */
  if (j->solver == CMTJ_RK4) {
    /* runSingleLayerSolver / runMultiLayerSolver, :1668-1707: neighbours
     * frozen at the start of the step for every stage of every layer */
    vec3 copies[CMTJ_MAX_LAYERS + 2];
    copies[0] = zero;
    copies[nl + 1] = zero;
    for (int l = 0; l < nl; ++l) copies[l + 1] = Ls[l].mag;
    for (int l = 0; l < nl; ++l) {
      if (rk4_step(&Ls[l], t, dt, i, copies[l], copies[l + 2])) {
        st->code = CMTJ_ERR_NAN;
        st->nan_layer = l;
        return 1;
      }
    }
  } else if (j->solver == CMTJ_EULER_HEUN) {
    /* eulerHeunSolverStep, :1709-1754 */
    vec3 mPrime[CMTJ_MAX_LAYERS];
    for (int l = 0; l < nl; ++l) {
      olayer* L = &Ls[l];
      const vec3 dW = langevin_vector(L, j, t, i, (uint32_t)l);
      const vec3 bottom = (l == 0) ? zero : Ls[l - 1].mag;
      const vec3 top = (l == nl - 1) ? zero : Ls[l + 1].mag;
      const vec3 fn = llg(L, t, i, 0, L->mag, bottom, top);
      const vec3 gn = stochastic_torque(L, L->mag, dW);
      const vec3 mNext = vadd(L->mag, vmul(gn, sqrt(dt)));
      const vec3 gnp = stochastic_torque(L, mNext, dW);
      mPrime[l] = vadd(vadd(L->mag, vmul(fn, dt)), vmul(smul(0.5, vadd(gn, gnp)), sqrt(dt)));
    }
    for (int l = 0; l < nl; ++l) Ls[l].mag = vnormalize(mPrime[l]);
  } else if (j->solver == CMTJ_HEUN) {
    /* heunSolverStep, :1756-1815 */
    vec3 fn[CMTJ_MAX_LAYERS], gn[CMTJ_MAX_LAYERS], dW[CMTJ_MAX_LAYERS], mNext[CMTJ_MAX_LAYERS];
    for (int l = 0; l < nl; ++l) {
      olayer* L = &Ls[l];
      const vec3 bottom = (l == 0) ? zero : Ls[l - 1].mag;
      const vec3 top = (l == nl - 1) ? zero : Ls[l + 1].mag;
      fn[l] = llg(L, t, i, 0, L->mag, bottom, top);
      dW[l] = langevin_vector(L, j, t, i, (uint32_t)l);
      gn[l] = stochastic_torque(L, L->mag, dW[l]);
      mNext[l] = vadd(vadd(L->mag, vmul(fn[l], dt)), vmul(gn[l], sqrt(dt)));
    }
    for (int l = 0; l < nl; ++l) {
      olayer* L = &Ls[l];
      const vec3 bottom = (l == 0) ? zero : mNext[l - 1];
      const vec3 top = (l == nl - 1) ? zero : mNext[l + 1];
      const vec3 f2 = llg(L, t + dt, i, 1, mNext[l], bottom, top);
      const vec3 g2 = stochastic_torque(L, mNext[l], dW[l]);
      vec3 m = vadd(L->mag, smul(0.5 * dt, vadd(fn[l], f2)));
      m = vadd(m, vmul(smul(0.5, vadd(gn[l], g2)), sqrt(dt)));
      L->mag = vnormalize(m);
    }
  } else {
    st->code = CMTJ_ERR_INVALID;
    return 2;
  }
  return 0;
}

static void store_mag(const cmtj_junction* j, const olayer* Ls, double* mag_out) {
  for (int l = 0; l < j->n_layers; ++l) {
    mag_out[3 * l + 0] = Ls[l].mag.x;
    mag_out[3 * l + 1] = Ls[l].mag.y;
    mag_out[3 * l + 2] = Ls[l].mag.z;
  }
}

/* Layer::dormand_prince_step (core/junction.hpp:866-1002) with the
 * neighbours bottom/top (zero vectors at the stack ends, :1688-1691) and the
 * layer's step controller: the standard one, safety * (1/ratio)^(1/5)
 * (:962-965), or the PID one (:953-961), whose state advances on every
 * attempt, accepted or not.  Returns 1 if accepted; *h is replaced by the
 * proposed next step either way; *nan set on a NaN. */
static int dp_step(olayer* L, double time, double* h_io, vec3 bottom, vec3 top, int* nan) {
  const double h = *h_io;
  const vec3 m_t = L->mag;
  const vec3 k1 = vmul(llg(L, time, 0, 0, m_t, bottom, top), h);
  const vec3 k2 = vmul(llg(L, time + (1.0 / 5.0) * h, 0, 0, vadd(m_t, vmul(k1, 1.0 / 5.0)), bottom, top), h);
  const vec3 k3 = vmul(llg(L, time + (3.0 / 10.0) * h, 0, 0,
                           vadd(vadd(m_t, vmul(k1, 3.0 / 40.0)), vmul(k2, 9.0 / 40.0)), bottom, top), h);
  const vec3 k4 = vmul(llg(L, time + (4.0 / 5.0) * h, 0, 0,
                           vadd(vadd(vadd(m_t, vmul(k1, 44.0 / 45.0)), vmul(k2, -56.0 / 15.0)),
                                vmul(k3, 32.0 / 9.0)), bottom, top), h);
  const vec3 k5 = vmul(llg(L, time + (8.0 / 9.0) * h, 0, 0,
                           vadd(vadd(vadd(vadd(m_t, vmul(k1, 19372.0 / 6561.0)), vmul(k2, -25360.0 / 2187.0)),
                                     vmul(k3, 64448.0 / 6561.0)), vmul(k4, -212.0 / 729.0)), bottom, top), h);
  const vec3 k6 = vmul(llg(L, time + 1.0 * h, 0, 0,
                           vadd(vadd(vadd(vadd(vadd(m_t, vmul(k1, 9017.0 / 3168.0)), vmul(k2, -355.0 / 33.0)),
                                          vmul(k3, 46732.0 / 5247.0)), vmul(k4, 49.0 / 176.0)),
                                vmul(k5, -5103.0 / 18656.0)), bottom, top), h);
  const vec3 m5s = vadd(vadd(vadd(vadd(vadd(m_t, vmul(k1, 35.0 / 384.0)), vmul(k3, 500.0 / 1113.0)),
                                  vmul(k4, 125.0 / 192.0)), vmul(k5, -2187.0 / 6784.0)), vmul(k6, 11.0 / 84.0));
  const vec3 k7 = vmul(llg(L, time + 1.0 * h, 0, 0, m5s, bottom, top), h);
  vec3 m_t5 = m5s;
  const vec3 m_t4 = vadd(vadd(vadd(vadd(vadd(vadd(m_t, vmul(k1, 5179.0 / 57600.0)), vmul(k3, 7571.0 / 16695.0)),
                                        vmul(k4, 393.0 / 640.0)), vmul(k5, -92097.0 / 339200.0)),
                              vmul(k6, 187.0 / 2100.0)), vmul(k7, 1.0 / 40.0));
  const vec3 e = vsub(m_t5, m_t4);
  const double error_norm = sqrt(e.x * e.x + e.y * e.y + e.z * e.z);
  const double mlen = sqrt(m_t.x * m_t.x + m_t.y * m_t.y + m_t.z * m_t.z);
  const double tol = L->p->dp_abs_tol + L->p->dp_rel_tol * mlen;
  double error_ratio = error_norm / tol;
  if (tol < 1e-15 || isnan(error_ratio) || isinf(error_ratio)) error_ratio = 10.0;
  double factor;
  if (L->p->dp_use_pid) {
    const double e_n = log(error_ratio);
    L->pid_integral += e_n;
    factor = exp(fma(L->p->dp_kp, e_n, fma(L->p->dp_ki, L->pid_integral, L->p->dp_kd * (e_n - L->pid_prev))));
    L->pid_prev = e_n;
  } else {
    factor = L->p->dp_safety * pow(1.0 / error_ratio, 0.2);
  }
  if (isnan(factor) || isinf(factor)) factor = L->p->dp_min_factor;
  factor = fmin(factor, L->p->dp_max_factor);
  factor = fmax(factor, L->p->dp_min_factor);
  double new_h = h * factor;
  if (isnan(new_h) || isinf(new_h) || new_h <= 0) new_h = h * L->p->dp_min_factor;
  *h_io = new_h;
  if (error_ratio <= 1.0) {
    m_t5 = vnormalize(m_t5);
    L->mag = m_t5;
    if (isnan(L->mag.x)) *nan = 1;
    return 1;
  }
  return 0;
}

/* The adaptive branch of Junction::runSimulation (:1995-2044): the clock
 * advances by the step proposed for the next step; a log entry, labelled
 * with the new time, after each accepted step that reaches the next
 * multiple of the write frequency.  At most n_steps attempts.
 * Several layers (runMultiLayerSolver, :1682-1707): neighbours are the
 * magnetisations at the start of the attempt, the layers step in order and
 * each takes the step size the previous one proposed (dormand_prince_step
 * writes its proposal into the shared step), and a layer that accepts keeps
 * its new m even when another rejects.  After a rejection the reference sets
 * the step to the smallest Layer::hopt, which is never assigned and stays -1
 * (:259), so the attempt is forced through with a step of -1 s and its clock
 * runs backwards; the run stops there with CMTJ_ERR_DP_REJECTED. */
static void dp_run(const cmtj_junction* j, olayer* Ls, double* log, cmtj_status* st) {
  const double T = j->total_time, wf = j->write_freq, h0 = j->dt;
  const int nl = j->n_layers;
  const size_t E = j->n_entries;
  const size_t tch = (size_t)oracle_n_channels(nl, j->mr_mode, j->log_mask) - 1;
  const vec3 zero = v3(0, 0, 0);
  double t = 0.0, h = h0, next_write = 0.0;
  uint32_t it = 0, k = 0, accepted = 0;
  while (t < T) {
    if (it >= j->n_steps) {
      st->code = CMTJ_ERR_LIMIT;
      break;
    }
    ++it;
    if (t + h > T) h = T - t;
    int rejections = 0;
    if (isnan(h) || isinf(h) || h <= 0) {
      h = h0;
      ++rejections;
    }
    vec3 copies[CMTJ_MAX_LAYERS + 2];
    copies[0] = zero;
    copies[nl + 1] = zero;
    for (int l = 0; l < nl; ++l) copies[l + 1] = Ls[l].mag;
    int acc = 1, nan = 0;
    for (int l = 0; l < nl; ++l) {
      if (!dp_step(&Ls[l], t, &h, copies[l], copies[l + 2], &nan)) acc = 0;
      if (nan) {
        st->code = CMTJ_ERR_NAN;
        st->nan_layer = l;
        break;
      }
    }
    if (nan) break;
    if (!acc && nl > 1) {
      st->code = CMTJ_ERR_DP_REJECTED;
      break;
    }
    if (!acc && (h < 1e-15 || rejections >= 10)) acc = 1;
    if (acc) {
      t += h;
      ++accepted;
      if (t >= next_write) {
        if (k >= E) {
          st->code = CMTJ_ERR_LIMIT;
          break;
        }
        if (j->log_mask) {
          log_entry(j, Ls, log, k);
          log[tch * E + k] = t;
        }
        ++k;
        next_write = (double)((unsigned int)(t / wf) + 1u) * wf;
      }
    }
  }
  st->steps_done = accepted;
  st->entries_done = k;
}

int oracle_run(const cmtj_junction* j, const cmtj_layer* layers, const cmtj_constants* c,
               const double* tables, size_t n_table, double* log, double* mag_out, cmtj_status* st) {
  (void)n_table;
  olayer Ls[CMTJ_MAX_LAYERS];
  load_layers(j, layers, c, tables, Ls);
  st->code = CMTJ_OK;
  st->steps_done = 0;
  st->entries_done = 0;
  st->nan_layer = -1;
  if (j->solver == CMTJ_DORMAND_PRINCE) {
    dp_run(j, Ls, log, st);
    store_mag(j, Ls, mag_out);
    int pid = 0;
    for (int l = 0; l < j->n_layers; ++l) pid |= layers[l].dp_use_pid;
    if (pid) /* the layers' final controller state after the magnetisations */
      for (int l = 0; l < j->n_layers; ++l) {
        mag_out[3 * j->n_layers + 2 * l] = Ls[l].pid_integral;
        mag_out[3 * j->n_layers + 2 * l + 1] = Ls[l].pid_prev;
      }
    return 0;
  }
  uint32_t k = 0;
  for (uint32_t i = 0; i < j->n_steps; ++i) {
    if (junction_step(j, Ls, i, st)) break;
    st->steps_done = i + 1;
    if (i % j->write_every == 0) { /* :1978-1980 */
      log_entry(j, Ls, log, k);
      ++k;
      st->entries_done = k;
    }
  }
  store_mag(j, Ls, mag_out);
  return 0;
}

/* ---- electrically coupled stacks (core/stack.hpp) ----------------------- */

/* Stack<T>::runSimulationKCL (:372-410) and runSimulationNonKCL with
 * delayed or non-delayed coupling (:305-370) of SeriesStack / ParallelStack
 * (:414-480).  Member q
 * logs into log + js[q].log_offset: its junction channels, then the stack's
 * I_q and C_q_{q+1} (logStackData, :206-217).  Per step the coupling uses
 * the members' magnetisations at the start of the step (storeMagnetisations,
 * :65-76); each member's current driver is replaced for the step by the
 * local driver (executeJunctionStep, :78-91: setLayerCurrentDriver("all")),
 * then the member takes one step of its solver. */
int oracle_run_stack(const cmtj_stack* s, const cmtj_junction* js, const cmtj_layer* layers,
                     const cmtj_constants* c, const double* tables, double* log, double* mag_out,
                     cmtj_status* st) {
  const int n = s->n_junctions;
  if (n < 2 || n > CMTJ_MAX_STACK) return 1;
  cmtj_layer* own = (cmtj_layer*)malloc(sizeof(cmtj_layer) * CMTJ_MAX_LAYERS * (size_t)n);
  olayer* Ls = (olayer*)malloc(sizeof(olayer) * CMTJ_MAX_LAYERS * (size_t)n);
  if (!own || !Ls) {
    free(own);
    free(Ls);
    return 1;
  }
  for (int q = 0; q < n; ++q) {
    const cmtj_junction* j = &js[q];
    memcpy(&own[q * CMTJ_MAX_LAYERS], layers + j->layer_offset, sizeof(cmtj_layer) * (size_t)j->n_layers);
    load_layers(j, &own[q * CMTJ_MAX_LAYERS], c, tables, &Ls[q * CMTJ_MAX_LAYERS]);
    st[q].code = CMTJ_OK;
    st[q].steps_done = 0;
    st[q].entries_done = 0;
    st[q].nan_layer = -1;
  }
  const cmtj_junction* j0 = &js[0];
  double C[CMTJ_MAX_STACK], Ilog[CMTJ_MAX_STACK], Clog[CMTJ_MAX_STACK];
  uint32_t k = 0;
  int stop = 0;
  for (uint32_t i = 0; i < j0->n_steps && !stop; ++i) {
    const double t = i * j0->dt;
    /* storeMagnetisations and getEffectiveCouplingStrength (:428-436, :459-466) */
    double d[CMTJ_MAX_STACK];
    for (int q = 0; q < n; ++q) {
      const olayer* L = &Ls[q * CMTJ_MAX_LAYERS];
      const vec3 m = L[s->top[q]].mag;
      const vec3 p = s->two_layer ? L[s->bottom[q]].mag : v3(s->pol[0], s->pol[1], s->pol[2]);
      d[q] = vdot(m, p);
    }
    for (int q = 0; q + 1 < n; ++q) C[q] = s->coupling[q] * (s->series ? d[q] + d[q + 1] : d[q] - d[q + 1]);
    cmtj_driver local[CMTJ_MAX_STACK];
    if (s->kcl) {
      /* a callback current: one sample per step (getCurrentScalarValue(t), :376) */
      const double I0 = oracle_driver_value(&s->current, t, i, 0, 1, tables);
      double total = I0;
      for (int q = 0; q + 1 < n; ++q) {
        total += C[q] * I0;
        Clog[q] = C[q];
      }
      for (int q = 0; q < n; ++q) {
        memset(&local[q], 0, sizeof(cmtj_driver));
        local[q].type = CMTJ_CONSTANT; /* getConstantDriver(totalCurrent); phaseShift leaves it */
        local[q].constant = total;
      }
    } else {
      double eff = 1.0;
      for (int q = 0; q < n; ++q) {
        if (q > 0) {
          eff *= (1 + C[q - 1]);
          Clog[q - 1] = eff;
        }
        local[q] = s->current; /* currentDriver * effectiveCoupling (drivers.hpp:313-318) */
        local[q].constant = s->current.constant * eff;
        local[q].amplitude = s->current.amplitude * eff;
        local[q].phase += s->series ? s->phase_offset * q : s->phase_offset; /* phaseShift */
      }
    }
    double eff_nd = 1.0;
    for (int q = 0; q < n && !stop; ++q) {
      const cmtj_junction* j = &js[q];
      if (!s->kcl && !s->delayed && q > 0) {
        /* non-delayed coupling (:334-343): member q-1 has taken this step,
         * member q has not; two-layer members, p = the bottom layer */
        const olayer* A = &Ls[(q - 1) * CMTJ_MAX_LAYERS];
        const olayer* B = &Ls[q * CMTJ_MAX_LAYERS];
        const double da = vdot(A[s->top[q - 1]].mag, A[s->bottom[q - 1]].mag);
        const double db = vdot(B[s->top[q]].mag, B[s->bottom[q]].mag);
        eff_nd *= (1 + s->coupling[q - 1] * (s->series ? da + db : da - db));
        Clog[q - 1] = eff_nd;
        local[q] = s->current;
        local[q].constant = s->current.constant * eff_nd;
        local[q].amplitude = s->current.amplitude * eff_nd;
        local[q].phase += s->series ? s->phase_offset * q : s->phase_offset;
      }
      for (int l = 0; l < j->n_layers; ++l) own[q * CMTJ_MAX_LAYERS + l].drv[CMTJ_SLOT_CURRENT] = local[q];
      if (junction_step(j, &Ls[q * CMTJ_MAX_LAYERS], i, &st[q])) stop = 1;
      Ilog[q] = oracle_driver_value(&local[q], t, i, 0, 1, NULL);
    }
    if (stop) break;
    for (int q = 0; q < n; ++q) st[q].steps_done = i + 1;
    if (i % j0->write_every == 0) { /* logSimulationStep, :93-104 */
      for (int q = 0; q < n; ++q) {
        const cmtj_junction* j = &js[q];
        double* blk = log + j->log_offset;
        log_entry(j, &Ls[q * CMTJ_MAX_LAYERS], blk, k);
        const size_t E = j->n_entries;
        const size_t nch = (size_t)oracle_n_channels(j->n_layers, j->mr_mode, j->log_mask);
        blk[nch * E + k] = Ilog[q];
        blk[(nch + 1) * E + k] = (q + 1 < n) ? Clog[q] : 0.0;
        st[q].entries_done = k + 1;
      }
      ++k;
    }
  }
  for (int q = 0; q < n; ++q)
    store_mag(&js[q], &Ls[q * CMTJ_MAX_LAYERS], mag_out + (size_t)q * 3 * CMTJ_MAX_LAYERS);
  free(own);
  free(Ls);
  return 0;
}

/* ---- dipole-coupled groups (core/reservoir.hpp) ------------------------ */

/* computeDipoleInteraction, core/reservoir.hpp:79-96 (Kanao et al.): the
 * field at r1 of the moment of layer2 at r2.  Returns 1 when the points are
 * too close (the reference throws). */
static int dipole_kanao(vec3 r1, vec3 r2, const cmtj_layer* layer2, vec3 m1, double mu0, vec3* out) {
  const vec3 rij = vsub(r1, r2);
  if (vlength(rij) < 1e-10) return 1;
  const double r3 = pow(vlength(rij), 3);
  const double r5 = pow(vlength(rij), 5);
  const double V = layer2->thickness * layer2->cell_surface;
  const double prefactor = (layer2->Ms / mu0) * V;
  *out = smul(prefactor, vsub(vdiv(smul(3 * vdot(m1, rij), rij), r5), vdiv(m1, r3)));
  return 0;
}

/* computeDipoleInteractionNoumra, :106-119 (Nomura et al.), with the tensor
 * of getDipoleTensorFromRelPositions (:36-56) applied by
 * calculate_tensor_interaction (core/junction.hpp:64-72). */
static int dipole_nomura(vec3 r1, vec3 r2, const cmtj_layer* layer2, vec3 m1, double mu0, vec3* out) {
  const vec3 rij = vsub(r1, r2);
  if (vlength(rij) < 1e-14) return 1;
  const double r_mag = pow(vlength(rij), 2);
  const double mult = 3 / (4 * M_PI * pow(vlength(rij), 5));
  double T[9];
  const vec3 t0 = vmul(v3(pow(rij.x, 2) - (r_mag / 3), rij.x * rij.y, rij.x * rij.z), mult);
  const vec3 t1 = vmul(v3(rij.x * rij.y, pow(rij.y, 2) - (r_mag / 3), rij.y * rij.z), mult);
  const vec3 t2 = vmul(v3(rij.x * rij.z, rij.y * rij.z, pow(rij.z, 2) - (r_mag / 3)), mult);
  T[0] = t0.x, T[1] = t0.y, T[2] = t0.z;
  T[3] = t1.x, T[4] = t1.y, T[5] = t1.z;
  T[6] = t2.x, T[7] = t2.y, T[8] = t2.z;
  const double V = layer2->thickness * layer2->cell_surface;
  *out = tensor_interaction(m1, T, layer2->Ms * V, mu0);
  return 0;
}

/* GroupInteraction::runSimulation (core/reservoir.hpp:181-221): each step,
 * with the top layers' magnetisations at its start, element i's top layer
 * gets the reserved interaction field sum_{j != i} f(r_i, r_j) (j
 * ascending, stepFunctionalSolver :126-147, AxialDriver of constants), then
 * the elements step one after another; entries every write_every steps.
 * Same descriptors and log layout as cmtj_run_groups for one group; mag_out
 * [n][CMTJ_MAX_LAYERS][3], st [n].  Returns 0, 1 for a bad descriptor, 2
 * when two points are too close. */
int oracle_run_group(const cmtj_group* g, const cmtj_junction* js, const cmtj_layer* layers,
                     const cmtj_constants* c, const double* pair_terms, double* log, double* mag_out,
                     cmtj_status* st) {
  const int n = g->n_elements;
  if (n < 1 || n > CMTJ_MAX_GROUP) return 1;
  cmtj_layer* own = (cmtj_layer*)malloc(sizeof(cmtj_layer) * CMTJ_MAX_LAYERS * (size_t)n);
  olayer* Ls = (olayer*)malloc(sizeof(olayer) * CMTJ_MAX_LAYERS * (size_t)n);
  if (!own || !Ls) {
    free(own);
    free(Ls);
    return 1;
  }
  for (int q = 0; q < n; ++q) {
    const cmtj_junction* j = &js[q];
    memcpy(&own[q * CMTJ_MAX_LAYERS], layers + j->layer_offset, sizeof(cmtj_layer) * (size_t)j->n_layers);
    load_layers(j, &own[q * CMTJ_MAX_LAYERS], c, NULL, &Ls[q * CMTJ_MAX_LAYERS]);
    st[q].code = CMTJ_OK;
    st[q].steps_done = 0;
    st[q].entries_done = 0;
    st[q].nan_layer = -1;
  }
  const cmtj_junction* j0 = &js[0];
  uint32_t k = 0;
  int stop = 0, rc = 0;
  for (uint32_t i = 0; i < j0->n_steps && !stop; ++i) {
    vec3 H[CMTJ_MAX_GROUP];
    for (int a = 0; a < n && !rc; ++a) {
      const vec3 ra = vload(g->pos[a]);
      H[a] = v3(0, 0, 0);
      for (int b = 0; b < n && !rc; ++b) {
        if (a == b) continue;
        const cmtj_layer* lb = &own[b * CMTJ_MAX_LAYERS + g->top[b]];
        const vec3 mb = Ls[b * CMTJ_MAX_LAYERS + g->top[b]].mag;
        vec3 f = v3(0, 0, 0); /* nullDipoleInteraction, :63-68 */
        if (g->interaction == CMTJ_INTERACTION_KANAO)
          rc = dipole_kanao(ra, vload(g->pos[b]), lb, mb, c->mu0, &f);
        else if (g->interaction == CMTJ_INTERACTION_NOMURA)
          rc = dipole_nomura(ra, vload(g->pos[b]), lb, mb, c->mu0, &f);
        else if (g->interaction == CMTJ_INTERACTION_TENSOR) {
          /* an affine callable, f = T_ab m_b + b_ab; terms [b][12][a] */
          const double* p = pair_terms + (size_t)b * 12 * (size_t)n + (size_t)a;
          const size_t sN = (size_t)n;
          f.x = fma(p[0], mb.x, fma(p[sN], mb.y, fma(p[2 * sN], mb.z, p[9 * sN])));
          f.y = fma(p[3 * sN], mb.x, fma(p[4 * sN], mb.y, fma(p[5 * sN], mb.z, p[10 * sN])));
          f.z = fma(p[6 * sN], mb.x, fma(p[7 * sN], mb.y, fma(p[8 * sN], mb.z, p[11 * sN])));
        }
        H[a] = vadd(H[a], f);
      }
    }
    if (rc) {
      rc = 2;
      break;
    }
    for (int q = 0; q < n; ++q) {
      cmtj_driver* d = own[q * CMTJ_MAX_LAYERS + g->top[q]].drv + CMTJ_SLOT_HRES;
      const double hv[3] = {H[q].x, H[q].y, H[q].z};
      for (int s = 0; s < 3; ++s) {
        memset(&d[s], 0, sizeof(cmtj_driver));
        d[s].type = CMTJ_CONSTANT;
        d[s].constant = hv[s];
      }
    }
    for (int q = 0; q < n && !stop; ++q)
      if (junction_step(&js[q], &Ls[q * CMTJ_MAX_LAYERS], i, &st[q])) stop = 1;
    if (stop) break;
    for (int q = 0; q < n; ++q) st[q].steps_done = i + 1;
    if (i % j0->write_every == 0) { /* logLayerParams of every element, :215-218 */
      for (int q = 0; q < n; ++q) {
        log_entry(&js[q], &Ls[q * CMTJ_MAX_LAYERS], log + js[q].log_offset, k);
        st[q].entries_done = k + 1;
      }
      ++k;
    }
  }
  for (int q = 0; q < n; ++q)
    store_mag(&js[q], &Ls[q * CMTJ_MAX_LAYERS], mag_out + (size_t)q * 3 * CMTJ_MAX_LAYERS);
  free(own);
  free(Ls);
  return rc;
}

/* ---- threaded batch (CPU baseline) ------------------------------------ */

typedef struct {
  const cmtj_junction* js;
  size_t n;
  const cmtj_layer* layers;
  const cmtj_constants* c;
  const double* tables;
  size_t n_table;
  double* log;
  double* mag_out;
  cmtj_status* st;
  size_t next;
  pthread_mutex_t mu;
} batch_ctx;

static void* batch_worker(void* arg) {
  batch_ctx* b = (batch_ctx*)arg;
  for (;;) {
    pthread_mutex_lock(&b->mu);
    const size_t i = b->next++;
    pthread_mutex_unlock(&b->mu);
    if (i >= b->n) break;
    const cmtj_junction* j = &b->js[i];
    double* lg = b->log ? b->log + j->log_offset : NULL;
    double scratch_mag[3 * CMTJ_MAX_LAYERS];
    oracle_run(j, b->layers + j->layer_offset, b->c, b->tables, b->n_table, lg,
               b->mag_out ? b->mag_out + (size_t)i * 3 * CMTJ_MAX_LAYERS : scratch_mag, &b->st[i]);
  }
  return NULL;
}

int oracle_run_batch(const cmtj_junction* js, size_t n, const cmtj_layer* layers,
                     const cmtj_constants* c, const double* tables, size_t n_table, double* log,
                     double* mag_out, cmtj_status* st, int n_threads) {
  batch_ctx b;
  b.js = js;
  b.n = n;
  b.layers = layers;
  b.c = c;
  b.tables = tables;
  b.n_table = n_table;
  b.log = log;
  b.mag_out = mag_out;
  b.st = st;
  b.next = 0;
  pthread_mutex_init(&b.mu, NULL);
  if (n_threads < 1) n_threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)n_threads);
  for (int t = 0; t < n_threads; ++t) pthread_create(&th[t], NULL, batch_worker, &b);
  for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&b.mu);
  return 0;
}
