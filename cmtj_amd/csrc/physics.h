// physics.h -- device (and host) building blocks of the macrospin engine.
//
// Waveforms of the reference ScalarDriver (core/drivers.hpp:262-305), the
// effective-field terms and LLG(S) torques of Layer<T> (core/junction.hpp:
// 579-831), written for one junction per GPU lane: all per-layer parameters
// are pre-derived once per lane, constant driver values are folded into a
// per-layer base vector, and only time-dependent drivers are evaluated per
// stage.  R is the arithmetic type of the field/torque stages (double for the
// FP64 drop-in, float for the FP32 bulk mode); time, driver phases and the
// state m stay in double.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../../include/cmtj_amd.h"

#define CMTJ_HD __host__ __device__ __forceinline__

namespace cmtj {

// ---- feature bits (batch-uniform) -----------------------------------------
enum Feature : uint32_t {
  F_DEMAG_OFF = 1u << 0,  // off-diagonal demag entries
  F_DIPOLE = 1u << 1,     // dipole tensors to neighbours
  F_ANIS = 1u << 2,       // first-order anisotropy
  F_ANIS2 = 1u << 3,      // second-order anisotropy
  F_IEC = 1u << 4,        // bilinear IEC
  F_IEC2 = 1u << 5,       // biquadratic IEC
  F_IDMI = 1u << 6,       // interlayer DMI
  F_STT = 1u << 7,        // Slonczewski torque
  F_SOT = 1u << 8,        // spin-orbit torque
  F_SOT2 = 1u << 9,       // secondary SOT reference
  F_THERMAL = 1u << 10,   // stochastic thermal field
  F_DYN = 1u << 11,       // time-dependent drivers on non-field slots
  F_TABLE = 1u << 12,     // tabulated (callback) drivers present
  F_DYNH = 1u << 13,      // time-dependent drivers that feed field slots only
  F_ALL = 0x3FFFu,        // every feature bit
  // kernel-only bits (never set in a group's runtime flags):
  F_EXACT = 1u << 16,     // the group's features are exactly the kernel's: feature tests are compile-time
  F_KSTEP = 1u << 17,     // every time-dependent record is a step driver
  F_KSINE = 1u << 18,     // every time-dependent record is a sine driver
  F_KTRAP = 1u << 19,     // every time-dependent record is a trapezoid driver
};

template <class R>
struct V3 {
  R x, y, z;
};

template <class R>
CMTJ_HD V3<R> mk(R x, R y, R z) {
  V3<R> r;
  r.x = x;
  r.y = y;
  r.z = z;
  return r;
}
template <class R>
CMTJ_HD V3<R> operator+(V3<R> a, V3<R> b) {
  return mk<R>(a.x + b.x, a.y + b.y, a.z + b.z);
}
template <class R>
CMTJ_HD V3<R> operator-(V3<R> a, V3<R> b) {
  return mk<R>(a.x - b.x, a.y - b.y, a.z - b.z);
}
template <class R>
CMTJ_HD V3<R> operator*(V3<R> a, R s) {
  return mk<R>(a.x * s, a.y * s, a.z * s);
}
template <class R>
CMTJ_HD R dot(V3<R> a, V3<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
CMTJ_HD V3<R> cross(V3<R> a, V3<R> b) {
  return mk<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R, class S>
CMTJ_HD V3<R> cvt(V3<S> a) {
  return mk<R>((R)a.x, (R)a.y, (R)a.z);
}

// ---- driver waveforms -------------------------------------------------------

// Time-dependent part f(t) of a ScalarDriver, value = constant + f(t)
// (core/drivers.hpp:262-305).  Exact double arithmetic in the order of the
// reference.  CMTJ_CONSTANT and CMTJ_TABLE return 0 here.  The transcendental
// kinds live out of line so that a kernel carries one copy of their code.
static __host__ __device__ __noinline__ double wave_part_trans(int type, double A, double p1, double p2,
                                                         double t) {
  switch (type) {
    case CMTJ_SINE:  // p1 frequency, p2 phase
      return A * sin(2 * M_PI * t * p1 + p2);
    case CMTJ_POSINE:
      return fabs(A * sin(2 * M_PI * t * p1 + p2));
    case CMTJ_HALFSINE: {
      const double v = A * sin(2 * M_PI * t * p1 + p2);
      return (v <= 0) ? v : 0.0;
    }
    case CMTJ_GAUSSIMPULSE: {  // p1 t0, p2 sigma
      const double d = t - p1;
      return A * exp(-(d * d) / (2 * (p2 * p2)));
    }
    case CMTJ_GAUSSSTEP:
      return 0.5 * A * (1 + erf((t - p1) / (sqrt(2.0) * p2)));
    default:
      return 0.0;
  }
}

// trapezoidDriver (drivers.hpp:90-108): p1 start, p2 edge, p3 steady
// trapezoid with its ramp slope A / p2 given (computed once per run by the
// kernels whose records are all trapezoids; the same value as the division)
CMTJ_HD double wave_trapezoid_slope(double A, double slope, double p1, double p2, double p3, double t) {
  if (t < p1) return 0.0;
  if (t <= p1 + p2) return slope * (t - p1);
  if (t <= p1 + p2 + p3) return A;
  if (t <= p1 + 2 * p2 + p3) return A - slope * (t - (p1 + p2 + p3));
  return 0.0;
}
CMTJ_HD double wave_trapezoid(double A, double p1, double p2, double p3, double t) {
  return wave_trapezoid_slope(A, A / p2, p1, p2, p3, t);
}

CMTJ_HD double wave_part(int type, double A, double p1, double p2, double p3, double t) {
  switch (type) {
    case CMTJ_PULSE: {  // p1 period, p2 cycle (pulseTrain, drivers.hpp:79-88)
      const int n = (int)(t / p1);
      const double dT = p2 * p1;
      const double nT = n * p1;
      return (nT <= t && t <= (nT + dT)) ? A : 0.0;
    }
    case CMTJ_STEP:  // p1 start, p2 stop, both inclusive (stepUpdate, :73-78)
      return (t >= p1 && t <= p2) ? A : 0.0;
    case CMTJ_TRAPEZOID:
      return wave_trapezoid(A, p1, p2, p3, t);
    case CMTJ_SINE:
    case CMTJ_POSINE:
    case CMTJ_HALFSINE:
    case CMTJ_GAUSSIMPULSE:
    case CMTJ_GAUSSSTEP:
      return wave_part_trans(type, A, p1, p2, t);
    default:
      return 0.0;
  }
}

// FP32-bulk variant: edges and phases in double; the sine phase is reduced
// to [-pi, pi] in double before the binary32 sine.
static __device__ __noinline__ float wave_part_f_trans(int type, double A, double p1, double p2, double t) {
  switch (type) {
    case CMTJ_SINE:
    case CMTJ_POSINE:
    case CMTJ_HALFSINE: {
      double x = 2 * M_PI * t * p1 + p2;
      x = x - (2 * M_PI) * rint(x * (1.0 / (2 * M_PI)));
      const float v = (float)A * sinf((float)x);
      if (type == CMTJ_SINE) return v;
      if (type == CMTJ_POSINE) return fabsf(v);
      return (v <= 0.f) ? v : 0.f;
    }
    case CMTJ_GAUSSIMPULSE: {
      const double d = t - p1;
      return (float)A * expf((float)(-(d * d) / (2 * (p2 * p2))));
    }
    case CMTJ_GAUSSSTEP:
      return 0.5f * (float)A * (1.f + erff((float)((t - p1) / (sqrt(2.0) * p2))));
    default:
      return 0.f;
  }
}

__device__ __forceinline__ float wave_part_f(int type, double A, double p1, double p2, double p3,
                                             double t) {
  switch (type) {
    case CMTJ_SINE:
    case CMTJ_POSINE:
    case CMTJ_HALFSINE:
    case CMTJ_GAUSSIMPULSE:
    case CMTJ_GAUSSSTEP:
      return wave_part_f_trans(type, A, p1, p2, t);
    default:
      return (float)wave_part(type, A, p1, p2, p3, t);
  }
}

// Parameters of the time-dependent part of a driver, by kind.
CMTJ_HD void wave_params(const cmtj_driver& d, double& A, double& p1, double& p2, double& p3) {
  A = d.amplitude;
  p1 = p2 = p3 = 0.0;
  switch (d.type) {
    case CMTJ_PULSE:
      p1 = d.period;
      p2 = d.cycle;
      break;
    case CMTJ_SINE:
    case CMTJ_POSINE:
    case CMTJ_HALFSINE:
      p1 = d.frequency;
      p2 = d.phase;
      break;
    case CMTJ_STEP:
      p1 = d.t_start;
      p2 = d.t_stop;
      break;
    case CMTJ_TRAPEZOID:
      p1 = d.t_start;
      p2 = d.edge_time;
      p3 = d.steady_time;
      break;
    case CMTJ_GAUSSIMPULSE:
    case CMTJ_GAUSSSTEP:
      p1 = d.t_start;
      p2 = d.edge_time;
      break;
    default:
      break;
  }
}

// Full driver value (host API and tests): constant + f(t).
CMTJ_HD double driver_value(const cmtj_driver& d, double t) {
  double A, p1, p2, p3;
  wave_params(d, A, p1, p2, p3);
  return d.constant + wave_part(d.type, A, p1, p2, p3, t);
}

// ---- Philox4x32-10 counter-based generator ---------------------------------

CMTJ_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                           uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// Three N(0,1) draws of (step, layer) in the junction's stream: Box-Muller
// over u = (x + 0.5) 2^-32, counter {step, layer, stream}, key seed.
template <class R>
__device__ __forceinline__ V3<R> thermal_normals(uint64_t seed, uint64_t stream, uint32_t step,
                                                 uint32_t layer);

template <>
__device__ __forceinline__ V3<double> thermal_normals<double>(uint64_t seed, uint64_t stream,
                                                              uint32_t step, uint32_t layer) {
  uint32_t o[4];
  philox4x32_10(step, layer, (uint32_t)stream, (uint32_t)(stream >> 32), (uint32_t)seed,
                (uint32_t)(seed >> 32), o);
  const double s = 2.3283064365386963e-10;
  const double u0 = ((double)o[0] + 0.5) * s, u1 = ((double)o[1] + 0.5) * s;
  const double u2 = ((double)o[2] + 0.5) * s, u3 = ((double)o[3] + 0.5) * s;
  const double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
  return mk<double>(r0 * cos(2.0 * M_PI * u1), r0 * sin(2.0 * M_PI * u1),
                    r1 * cos(2.0 * M_PI * u3));
}

template <>
__device__ __forceinline__ V3<float> thermal_normals<float>(uint64_t seed, uint64_t stream,
                                                            uint32_t step, uint32_t layer) {
  uint32_t o[4];
  philox4x32_10(step, layer, (uint32_t)stream, (uint32_t)(stream >> 32), (uint32_t)seed,
                (uint32_t)(seed >> 32), o);
  const float s = 2.3283064365386963e-10f;
  const float u0 = ((float)o[0] + 0.5f) * s, u1 = ((float)o[1] + 0.5f) * s;
  const float u2 = ((float)o[2] + 0.5f) * s, u3 = ((float)o[3] + 0.5f) * s;
  // (float)o + 0.5 rounds to 2^32 for the top values; clamp below 1
  const float one_m = 0.99999994f;
  const float r0 = sqrtf(-2.0f * logf(fminf(u0, one_m)));
  const float r1 = sqrtf(-2.0f * logf(fminf(u2, one_m)));
  float s1, c1, s3, c3;
  sincospif(2.0f * u1, &s1, &c1);
  sincospif(2.0f * u3, &s3, &c3);
  return mk<float>(r0 * c1, r0 * s1, r1 * c3);
}

}  // namespace cmtj
