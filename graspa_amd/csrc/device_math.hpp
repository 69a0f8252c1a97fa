// device_math.hpp -- pair terms, minimum image, counter-based random numbers
// and wave64 reductions shared by the engine's kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fast_math_coeffs.hpp"

namespace graspa {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): a counter-based generator,
// so every insertion owns a stream addressed by (seed, stream, index, draw).
// ---------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
  }
}

// Uniform double in (0,1): 53 bits from two output words of one Philox block.
// Draws 2j and 2j+1 share the block with counter word 2 = j.
__host__ __device__ inline double philox_uniform(uint64_t seed, uint32_t stream, uint64_t index, uint32_t draw)
{
  uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), draw >> 1, stream};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint32_t a = (draw & 1u) ? c[2] : c[0];
  const uint32_t b = (draw & 1u) ? c[3] : c[1];
  const uint64_t x = (((uint64_t)a << 32) | b) >> 11;
  return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Minimum image in FP64, rounding half away from zero (gRASPA maths.cuh:427-450).
// ---------------------------------------------------------------------------
__device__ inline void pbc_fp64(double& x, double& y, double& z, const double* cell, const double* inv, int cubic)
{
  if (cubic) {
    x -= (double)(int)(x * inv[0] + ((x >= 0.0) ? 0.5 : -0.5)) * cell[0];
    y -= (double)(int)(y * inv[4] + ((y >= 0.0) ? 0.5 : -0.5)) * cell[4];
    z -= (double)(int)(z * inv[8] + ((z >= 0.0) ? 0.5 : -0.5)) * cell[8];
  } else {
    double s0 = inv[0] * x + inv[3] * y + inv[6] * z;
    double s1 = inv[1] * x + inv[4] * y + inv[7] * z;
    double s2 = inv[2] * x + inv[5] * y + inv[8] * z;
    s0 -= (double)(int)(s0 + ((s0 >= 0.0) ? 0.5 : -0.5));
    s1 -= (double)(int)(s1 + ((s1 >= 0.0) ? 0.5 : -0.5));
    s2 -= (double)(int)(s2 + ((s2 >= 0.0) ? 0.5 : -0.5));
    x = cell[0] * s0 + cell[3] * s1 + cell[6] * s2;
    y = cell[1] * s0 + cell[4] * s1 + cell[7] * s2;
    z = cell[2] * s0 + cell[5] * s1 + cell[8] * s2;
  }
}

// Minimum image for the fast form: nearest image by rint (ties to even; the
// reference rounds ties away from zero, which differs only at exactly half a
// cell, where the two images are equidistant).
template <bool CUBIC>
__device__ inline void pbc_rint(double& x, double& y, double& z, const double* cell, const double* inv)
{
  if (CUBIC) {
    x = fma(-rint(x * inv[0]), cell[0], x);
    y = fma(-rint(y * inv[4]), cell[4], y);
    z = fma(-rint(z * inv[8]), cell[8], z);
  } else {
    double s0 = inv[0] * x + inv[3] * y + inv[6] * z;
    double s1 = inv[1] * x + inv[4] * y + inv[7] * z;
    double s2 = inv[2] * x + inv[5] * y + inv[8] * z;
    s0 -= rint(s0);
    s1 -= rint(s1);
    s2 -= rint(s2);
    x = cell[0] * s0 + cell[3] * s1 + cell[6] * s2;
    y = cell[1] * s0 + cell[4] * s1 + cell[7] * s2;
    z = cell[2] * s0 + cell[5] * s1 + cell[8] * s2;
  }
}

// 12-6 LJ at full coupling (gRASPA maths.cuh:452-494 with scaling = 1):
// ff = {4 eps, sigma^2, shift}.
__device__ inline double lj_fp64(double rr, double a4e, double s2, double shift)
{
  const double temp = rr / s2;
  const double temp3 = temp * temp * temp;
  const double rri3 = 1.0 / (temp3 + 0.0);
  return a4e * (rri3 * (rri3 - 1.0)) - shift;
}

// ---------------------------------------------------------------------------
// Mixed precision: positions are 32-bit fixed-point cell fractions, so the
// difference of two positions wraps to the minimum image exactly in integer
// arithmetic (fraction in [-1/2, 1/2)).
//
// Fast ("mixed") form of one pair.  On gfx950 FP64 FMA issues at the same
// lane rate as unpacked FP32, so the arithmetic stays FP64 and only the seeds
// of the reciprocals come from the FP32 hardware: r^2 in FP64; 1/r^2 and 1/r
// seeded by v_rcp_f32 / v_rsq_f32 and refined by one FP64 Newton step
// (relative error ~1e-14); erfc(alpha r)/r from a per-box FP64 polynomial
// table (coul_table) or the FP64 exp * polynomial erfc_fast below.  An FP32
// LJ 12th power would amplify a 1e-7 error in r^2 twelve-fold, which this
// avoids.
__device__ inline double r2_fixed64(uint32_t ax, uint32_t ay, uint32_t az, uint32_t bx, uint32_t by, uint32_t bz,
                                    const double* h, int cubic)
{
  const double fx = (double)(int32_t)(ax - bx);
  const double fy = (double)(int32_t)(ay - by);
  const double fz = (double)(int32_t)(az - bz);
  if (cubic) {
    const double dx = fx * h[0], dy = fy * h[4], dz = fz * h[8];
    return dx * dx + dy * dy + dz * dz;
  }
  const double dx = h[0] * fx + h[3] * fy + h[6] * fz;
  const double dy = h[1] * fx + h[4] * fy + h[7] * fz;
  const double dz = h[2] * fx + h[5] * fy + h[8] * fz;
  return dx * dx + dy * dy + dz * dz;
}

__device__ inline double lj_mixed(double rr, double a4e, double s2, double shift)
{
  double y = (double)__builtin_amdgcn_rcpf((float)rr);   // v_rcp_f32 seed
  y = y * (2.0 - rr * y);
  const double sr = s2 * y;
  const double s3 = sr * sr * sr;
  return fma(a4e, fma(s3, s3, -s3), -shift);      // s3 (s3 - 1) in one rounding
}

// The same 12-6 term from c12 = 4 eps sigma^12 and c6 = 4 eps sigma^6:
// y^3 (c12 y^3 - c6) - shift with y = 1/r^2 (one multiply fewer than lj_mixed).
__device__ inline double lj_mixed_c(double rr, double c12, double c6, double shift)
{
  double y = (double)__builtin_amdgcn_rcpf((float)rr);   // v_rcp_f32 seed
  y = y * (2.0 - rr * y);
  const double y3 = y * y * y;
  return fma(y3, fma(c12, y3, -c6), -shift);
}

// exp(y) for -708 < y <= 0: Cody-Waite reduction y = n ln2 + r with a
// two-part ln 2, degree-11 Taylor polynomial on |r| <= ln2/2 (truncation
// 6e-15 relative), scaled by 2^n.
__device__ inline double exp_fast(double y)
{
  const double n = rint(y * fastmath::kLog2e);
  double r = fma(-n, fastmath::kLn2Hi, y);
  r = fma(-n, fastmath::kLn2Lo, r);
  double p = fastmath::kExpPoly[fastmath::kExpDeg];
#pragma unroll
  for (int k = fastmath::kExpDeg - 1; k >= 0; --k) p = fma(p, r, fastmath::kExpPoly[k]);
  return ldexp(p, (int)n);
}

// erfc(x) for x >= 0 to ~1e-13 relative: exp(-x^2) * t * F(t) with
// t = c / (c + x) (reciprocal seeded by v_rcp_f32, one FP64 Newton step) and
// F a fitted polynomial (fast_math_coeffs.hpp).  0 beyond x = 6, where
// erfc < 2.2e-17.
__device__ inline double erfc_fast(double x)
{
  const double d = fastmath::kErfcC + x;
  double y = (double)__builtin_amdgcn_rcpf((float)d);
  y = fma(y, fma(-d, y, 1.0), y);
  const double t = fastmath::kErfcC * y;
  const double u = fma(t, fastmath::kErfcUScale, fastmath::kErfcUOff);
  double F = fastmath::kErfcPoly[fastmath::kErfcDeg];
#pragma unroll
  for (int k = fastmath::kErfcDeg - 1; k >= 0; --k) F = fma(F, u, fastmath::kErfcPoly[k]);
  const double r = exp_fast(-x * x) * t * F;
  return (x < fastmath::kErfcXMax) ? r : 0.0;
}

// qq = prefactor * q_a * q_b; 1/r seeded by v_rsq_f32 and refined by one FP64
// Newton step (relative error ~2e-14).
__device__ inline double coulomb_mixed(double rr, double qq, double alpha)
{
  double y = (double)__builtin_amdgcn_rsqf((float)rr);   // v_rsq_f32 seed
  y = y * (1.5 - 0.5 * rr * y * y);
  return qq * erfc_fast(alpha * (rr * y)) * y;
}

// Squared distance in FP32 from fixed-point fractions, for the overlap
// screen (relative error below 1e-6; the screen thresholds carry a 1e-5
// margin).  h = cell * 2^-32.
template <bool CUBIC>
__device__ inline float r2_fixed32(uint32_t ax, uint32_t ay, uint32_t az, uint32_t bx, uint32_t by, uint32_t bz,
                                   const float* h)
{
  const float fx = (float)(int32_t)(ax - bx);
  const float fy = (float)(int32_t)(ay - by);
  const float fz = (float)(int32_t)(az - bz);
  if (CUBIC) {
    const float dx = fx * h[0], dy = fy * h[4], dz = fz * h[8];
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
  }
  const float dx = fmaf(h[6], fz, fmaf(h[3], fy, h[0] * fx));
  const float dy = fmaf(h[7], fz, fmaf(h[4], fy, h[1] * fx));
  const float dz = fmaf(h[8], fz, fmaf(h[5], fy, h[2] * fx));
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// ---------------------------------------------------------------------------
// wave64 reductions
// ---------------------------------------------------------------------------
// Value of v in lane l (l uniform across the wave).
__device__ inline double readlane_d(double v, int l)
{
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// v of the lane the DPP control CTRL names (a lane of a row ROW_MASK leaves
// out reads 0), both 32-bit halves moved by the data-parallel primitives.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline double dpp_d(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over the 64 lanes (every lane active), in every lane: pairs and quads
// by quad permutes, rows of 16 by rotations by 4 and 8, then the row sums
// carried into the last row by the row broadcasts of lanes 15 and 31, and
// lane 63 read back.
__device__ inline double wave_sum(double v)
{
  v += dpp_d<0xB1>(v);         // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);         // quad_perm [2,3,0,1]
  v += dpp_d<0x124>(v);        // row_ror:4
  v += dpp_d<0x128>(v);        // row_ror:8
  v += dpp_d<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
  v += dpp_d<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
  return readlane_d(v, 63);
}
__device__ inline uint32_t wave_or(uint32_t v)
{
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v |= __shfl_xor(v, m, 64);
  return v;
}
// Maximum over the 64 lanes (every lane active), in every lane; the masked
// broadcast steps fold in -inf, which leaves a maximum unchanged.
__device__ inline double wave_max(double v)
{
  constexpr double ninf = -__builtin_huge_val();
  v = fmax(v, dpp_d<0xB1>(v));
  v = fmax(v, dpp_d<0x4E>(v));
  v = fmax(v, dpp_d<0x124>(v));
  v = fmax(v, dpp_d<0x128>(v));
  const double b15 = dpp_d<0x142, 0xa>(v);
  v = fmax(v, ((threadIdx.x >> 4) & 1) ? b15 : ninf);
  const double b31 = dpp_d<0x143, 0xc>(v);
  v = fmax(v, ((threadIdx.x >> 5) & 1) ? b31 : ninf);
  return readlane_d(v, 63);
}
// Orders this wave's LDS writes before its later LDS reads by other lanes
// (a wave executes its LDS operations in order; this keeps the compiler from
// moving them across).
__device__ inline void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace graspa
