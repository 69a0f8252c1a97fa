// move_math.hpp -- rigid-body rotations, e^{ik.r} axis tables and the
// real-space Coulomb table lookup shared by the Widom and batched-GCMC kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "engine.hpp"

namespace graspa {

__device__ inline double2 cmul2(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Rotate_Quaternions (mc_utilities.h:423-457): uniform random rotation from
// three uniforms (Shoemake's quaternion).
__device__ inline void quat_rotate(double v[3], double u, double vv, double w)
{
  const double pi = 3.14159265358979323846;
  const double q0 = sqrt(1 - u) * sin(2 * pi * vv);
  const double q1 = sqrt(1 - u) * cos(2 * pi * vv);
  const double q2 = sqrt(u) * sin(2 * pi * w);
  const double q3 = sqrt(u) * cos(2 * pi * w);
  const double a01 = q0 * q1, a02 = q0 * q2, a03 = q0 * q3;
  const double a11 = q1 * q1, a12 = q1 * q2, a13 = q1 * q3;
  const double a22 = q2 * q2, a23 = q2 * q3, a33 = q3 * q3;
  const double r0 = 1.0 - 2.0 * (a22 + a33), r1 = 2.0 * (a12 - a03), r2 = 2.0 * (a13 + a02);
  const double r3 = 2.0 * (a12 + a03), r4 = 1.0 - 2.0 * (a11 + a33), r5 = 2.0 * (a23 - a01);
  const double r6 = 2.0 * (a13 - a02), r7 = 2.0 * (a23 + a01), r8 = 1.0 - 2.0 * (a11 + a22);
  const double x = v[0] * r0 + v[1] * r1 + v[2] * r2;
  const double y = v[0] * r3 + v[1] * r4 + v[2] * r5;
  const double z = v[0] * r6 + v[1] * r7 + v[2] * r8;
  v[0] = x; v[1] = y; v[2] = z;
}

// RotationAroundAxis (mc_utilities.h:459-483): rotation of p by theta about
// the unit axis (ax, ay, az).
__device__ inline void rotate_axis(double p[3], double theta, double ax, double ay, double az)
{
  double s, c;
  sincos(theta, &s, &c);
  const double w = 1.0 - c;
  const double r0 = ax * ax * w + c, r1 = ax * ay * w + az * s, r2 = ax * az * w - ay * s;
  const double r3 = ax * ay * w - az * s, r4 = ay * ay * w + c, r5 = ay * az * w + ax * s;
  const double r6 = ax * az * w + ay * s, r7 = ay * az * w - ax * s, r8 = az * az * w + c;
  const double x = p[0] * r0 + p[1] * r1 + p[2] * r2;
  const double y = p[0] * r3 + p[1] * r4 + p[2] * r5;
  const double z = p[0] * r6 + p[1] * r7 + p[2] * r8;
  p[0] = x; p[1] = y; p[2] = z;
}

// e^{i n s} along one cell axis for n = 0..kmax, s = 2 pi (inverse cell . r)
// on that axis, by the complex recurrence of the reference
// (Ewald_Energy_Functions.h:96-104).
__device__ inline void axis_table(const BoxConst& bc, double px, double py, double pz, int axis, double2* out)
{
  const double* a = bc.inv;
  double s;
  if (axis == 0) s = a[0] * px + a[3] * py + a[6] * pz;
  else if (axis == 1) s = a[1] * px + a[4] * py + a[7] * pz;
  else s = a[2] * px + a[5] * py + a[8] * pz;
  s *= 2 * 3.14159265358979323846;
  const int km = bc.kmax[axis];
  out[0] = make_double2(1.0, 0.0);
  double sn, cs;
  sincos(s, &sn, &cs);
  const double2 e1 = make_double2(cs, sn);
  if (km >= 1) out[1] = e1;
  // the running power stays in registers (out may be LDS: no store-to-load
  // round trip per step)
  double2 ek = e1;
  for (int k = 2; k <= km; ++k) {
    ek = cmul2(ek, e1);
    out[k] = ek;
  }
}

// erfc(alpha r)/r from the per-box table (coulomb_table.cpp): the interval
// is read from the exponent and top mantissa bits of r^2 as FP32, the value is
// a degree-kCoulTabDeg polynomial in r^2 - s_i (Horner).  Clamped indices only
// occur for r^2 below 2^-7, i.e. pairs the overlap rule already rejects.
__device__ inline double coul_table(const double2* __restrict__ ct, const double rr)
{
  constexpr int kShift = 23 - kCoulTabMBits;
  const uint32_t bits = __float_as_uint((float)rr);
  // r^2 below the first octave (overlaps) reads row 0.  Every caller
  // evaluates pairs inside the Coulomb cutoff only, and the table's last
  // octave ends above cut^2 by more than FP32 rounding of r^2, so the index
  // stays below the row count without a clamp
  // (one saturating subtract, v_sub_u32 with clamp)
  const unsigned idx = __builtin_elementwise_sub_sat(bits >> kShift, (unsigned)((127 + kCoulTabELo) << kCoulTabMBits));
  const double si = (double)__uint_as_float(bits & ~((1u << kShift) - 1u));
  const double d = rr - si;
  // the LDS copy's row stride; idx < 2^24, and the 24-bit multiply issues at
  // full rate where a 32-bit one takes four cycles
  const double2* c = ct + __umul24((unsigned)idx, (unsigned)kCoulLdsRow);
  double2 cc[kCoulTabStride / 2];
#pragma unroll
  for (int k = 0; k < kCoulTabStride / 2; ++k) cc[k] = c[k];
  double g = cc[kCoulTabStride / 2 - 1].y;
  g = fma(g, d, cc[kCoulTabStride / 2 - 1].x);
#pragma unroll
  for (int k = kCoulTabStride / 2 - 2; k >= 0; --k) {
    g = fma(g, d, cc[k].y);
    g = fma(g, d, cc[k].x);
  }
  return g;
}

// Cell fractions of a position as 32-bit fixed point (the screen's and the
// cell grid's coordinates).
__device__ inline uint4 to_fixed4(const BoxConst& bc, const double p[3])
{
  double s[3];
  s[0] = bc.inv[0] * p[0] + bc.inv[3] * p[1] + bc.inv[6] * p[2];
  s[1] = bc.inv[1] * p[0] + bc.inv[4] * p[1] + bc.inv[7] * p[2];
  s[2] = bc.inv[2] * p[0] + bc.inv[5] * p[1] + bc.inv[8] * p[2];
  uint32_t f[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double fr = s[d] - floor(s[d]);
    f[d] = (uint32_t)(int64_t)(fr * 4294967296.0);
  }
  return make_uint4(f[0], f[1], f[2], 0u);
}

// Squared-distance threshold of the overlap screen for one (atom type, site
// type) pair: E_LJ(r^2) = 4eps (x^2 - x) - shift with x = (sigma^2/r^2)^3
// exceeds OverlapCriteria exactly when x > (1 + sqrt(1 + 4c))/2,
// c = (OverlapCriteria + shift)/(4 eps) > 0, i.e. r^2 < sigma^2 / x^(1/3); the
// reference also flags r^2 < 0.01; both only inside the VDW cutoff
// (overlap_r2).  The FP32 value is lowered by 1e-5 relative so that a screen
// hit is a certain overlap in both precision forms; pairs near the threshold
// are left to the exact evaluation.
__device__ inline float screen_threshold(const double* f, double overlap, double cut2)
{
  return (float)(overlap_r2(f[0], f[1], f[2], overlap, cut2) * (1.0 - 1.0e-5));
}

// ---------------------------------------------------------------------------
// Flexible linear chains.  No counterpart in the reference, which grows rigid
// molecules only (read_data.cpp:2104-2116).  Bead i >= 1 sits at bond length
// from bead i - 1; the angle (i-2, i-1, i) has U = k/2 (theta - theta0)^2;
// the dihedral (i-3 .. i) has U = t0 + t1 (1 + cos phi) + t2 (1 - cos 2 phi)
// + t3 (1 + cos 3 phi), phi = pi trans.  A species' model is kChainDoubles
// doubles (engine.hpp).
// ---------------------------------------------------------------------------
// Draws of bead i, trial j of a move: the block at kFlexDraw + ((i - 1) * 32
// + j) * kFlexBlock.  Bead 1: draws 0, 1; bend attempt a: 3a, 3a + 1 (a
// Box-Muller normal), 3a + 2 (acceptance), a < kFlexBendTries; dihedral
// attempt a: 512 + 2a (angle), 513 + 2a (acceptance), a < kFlexTorsTries;
// bead 2's azimuth: kFlexBlock - 1.  Step i's selection uniform: draw
// kFlexDraw - 8 + i.
constexpr uint32_t kFlexDraw = 4096;
constexpr uint32_t kFlexBlock = 1024;
constexpr int kFlexBendTries = 160;
constexpr int kFlexTorsTries = 250;

__device__ inline double flex_u(uint64_t seed, uint32_t stream, uint64_t move, int i, int j, int d)
{
  return philox_uniform(seed, stream, move, kFlexDraw + (uint32_t)((i - 1) * 32 + j) * kFlexBlock + (uint32_t)d);
}

// Trial position of bead i (1 <= i < kChainBeads) given beads 0 .. i-1 in
// prev: bead 1 uniform on the bond sphere; the bend angle from
// sin(theta) exp(-beta U_bend) (normal proposal carrying the Boltzmann factor,
// accepted with probability sin(theta)); the dihedral of bead 3 from
// exp(-beta U_tors) (uniform proposal accepted with exp(-beta (U - U_lb)),
// U_lb = t0 + min(0, 2 t1) + min(0, 2 t2) + min(0, 2 t3)), bead 2's azimuth
// uniform.  Position: C + bond (-cos th bc + sin th (cos phi m + sin phi n)),
// bc = unit(C - B), n = unit((B - A) x bc) (bead 2: the unit normal built
// from the axis least aligned with bc), m = n x bc.  A sampler that runs out
// of attempts returns theta0 or pi.
__device__ __noinline__ void flex_trial(const double (*prev)[3], const int i, const double* __restrict__ chain,
                                  const double beta, const uint64_t seed, const uint32_t stream, const uint64_t move,
                                  const int j, double out[3])
{
  const double pi = 3.14159265358979323846;
  const double* cb = chain + kChainPer * i;
  const double bl = cb[0];
  const double* C = prev[i - 1];
  if (i == 1) {
    const double z = 2.0 * flex_u(seed, stream, move, i, j, 0) - 1.0;
    const double ph = 2.0 * pi * flex_u(seed, stream, move, i, j, 1);
    const double r = sqrt(fmax(0.0, 1.0 - z * z));
    out[0] = C[0] + bl * r * cos(ph);
    out[1] = C[1] + bl * r * sin(ph);
    out[2] = C[2] + bl * z;
    return;
  }
  const double* Bp = prev[i - 2];
  double bc[3] = {C[0] - Bp[0], C[1] - Bp[1], C[2] - Bp[2]};
  double r = sqrt(bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2]);
  bc[0] /= r; bc[1] /= r; bc[2] /= r;
  double n[3];
  if (i == 2) {
    double ax[3] = {0.0, 0.0, 0.0};
    const double fx = fabs(bc[0]), fy = fabs(bc[1]), fz = fabs(bc[2]);
    ax[(fx <= fy && fx <= fz) ? 0 : (fy <= fz ? 1 : 2)] = 1.0;
    n[0] = bc[1] * ax[2] - bc[2] * ax[1];
    n[1] = bc[2] * ax[0] - bc[0] * ax[2];
    n[2] = bc[0] * ax[1] - bc[1] * ax[0];
  } else {
    const double* A = prev[i - 3];
    const double ab[3] = {Bp[0] - A[0], Bp[1] - A[1], Bp[2] - A[2]};
    n[0] = ab[1] * bc[2] - ab[2] * bc[1];
    n[1] = ab[2] * bc[0] - ab[0] * bc[2];
    n[2] = ab[0] * bc[1] - ab[1] * bc[0];
  }
  r = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  n[0] /= r; n[1] /= r; n[2] /= r;
  const double m[3] = {n[1] * bc[2] - n[2] * bc[1], n[2] * bc[0] - n[0] * bc[2], n[0] * bc[1] - n[1] * bc[0]};
  // bend angle
  const double th0 = cb[1];
  const double sd = 1.0 / sqrt(beta * cb[2]);
  double th = th0;
  for (int a = 0; a < kFlexBendTries; ++a) {
    const double u1 = flex_u(seed, stream, move, i, j, 3 * a);
    const double u2 = flex_u(seed, stream, move, i, j, 3 * a + 1);
    const double ua = flex_u(seed, stream, move, i, j, 3 * a + 2);
    const double t = th0 + sd * sqrt(-2.0 * log(u1)) * cos(2.0 * pi * u2);
    if (t > 0.0 && t < pi && ua < sin(t)) {
      th = t;
      break;
    }
  }
  double ph;
  if (i == 2) {
    ph = 2.0 * pi * flex_u(seed, stream, move, i, j, kFlexBlock - 1);
  } else {
    const double* t = cb + 3;
    const double lb = t[0] + fmin(0.0, 2.0 * t[1]) + fmin(0.0, 2.0 * t[2]) + fmin(0.0, 2.0 * t[3]);
    ph = pi;
    for (int a = 0; a < kFlexTorsTries; ++a) {
      const double p = 2.0 * pi * flex_u(seed, stream, move, i, j, 512 + 2 * a);
      const double ua = flex_u(seed, stream, move, i, j, 513 + 2 * a);
      const double u = t[0] + t[1] * (1.0 + cos(p)) + t[2] * (1.0 - cos(2.0 * p)) + t[3] * (1.0 + cos(3.0 * p));
      if (ua < exp(-beta * (u - lb))) {
        ph = p;
        break;
      }
    }
  }
  const double d0 = -bl * cos(th), d1 = bl * sin(th) * cos(ph), d2 = bl * sin(th) * sin(ph);
#pragma unroll
  for (int k = 0; k < 3; ++k) out[k] = C[k] + d0 * bc[k] + d1 * m[k] + d2 * n[k];
}

}  // namespace graspa
