// widom_kernel.hip -- batched Widom / CBMC test insertions in one persistent
// launch (the throughput path of the GCMC energy engine on gfx950).
//
// One wavefront carries one insertion at a time; waves take insertions from
// an atomic counter, so the work balances itself and a result depends only on
// the insertion's index (counter-based random numbers).  A workgroup stages
// the overlap screen's cell table in LDS once (when it fits); after that the
// insertion loop has no workgroup barrier.  One insertion:
//   1. all its random draws at once, one Philox4x32-10 block per lane;
//   2. K first-bead trial positions (per-wave LDS);
//   3. overlap screen: FP32 squared distances from fixed-point coordinates
//      against per-(site type, atom type) thresholds r_ov^2 below which the
//      pair energy certainly exceeds OverlapCriteria (or r^2 < 0.01), over the
//      atoms of the 27 grid cells around the trial site (cells at least r_ov
//      wide, so this finds every overlapping atom); a trial with a hit is
//      rejected without computing its energies;
//   4. full pair energies (LJ + real-space Coulomb, FP64 accumulation) of the
//      trials the screen left, which also settles their overlap flags exactly;
//   5. log-sum-exp selection (one uniform), W_fb;
//   6. O orientations of the template about the selected first bead, the same
//      screen + evaluation over O x (M-1) sites, selection, W_ch;
//   7. Fourier-space energy of the selected molecule in FP64 against the
//      resident structure factors (active k-vectors only, one lane per k);
//   8. W = W_fb W_ch exp(-beta (dE_ewald + dE_tail)); a per-insertion record.
// Per-insertion records are summed in index order by two small kernels, so the
// batch sums are deterministic.
// The semantics are Insertion_Body (mc_swap_utilities.h:3-133) with
// Widom_Move_FirstBead_PARTIAL / Widom_Move_Chain_PARTIAL (mc_widom.h:305-614);
// a trial rejected by the screen is one the reference flags (VDW_Coulomb.cu:
// 1299-1300) and leaves out of the selection (Host_sum_Widom_HGGG_SEPARATE,
// mc_widom.h:41-87).
#include <cstring>

#include "device_math.hpp"
#include "engine.hpp"

namespace graspa {

constexpr int kMaxTrials = 32;   // K and O
constexpr int kMaxMol = 8;       // sites per molecule
constexpr int kMaxPts = 64;      // trial sites of one pass: K, or O x (M-1)
constexpr int kMaxDraws = 3 * kMaxTrials + 1 + 3 * kMaxTrials + 1;
constexpr int kSums = 6;         // W, W^2, W dE, n, n first bead placed, n complete
constexpr int kRedBlock = 256;
constexpr int kRedPer = 8;       // records per thread in the first reduction pass
// Waves per workgroup at most: the FP64 erfc/exp coefficients of the pair term
// keep ~190 VGPRs live, i.e. 2 waves per SIMD; 8 waves fill a CU with one workgroup.
constexpr int kMaxWaves = 8;

// Per-wave LDS workspace.  The Ewald phase reuses the draw/trial area for the
// e^{ik.r} axis tables of the selected molecule.
struct WaveLds {
  double ud[kMaxDraws];          // this insertion's uniforms
  double pos[kMaxPts][3];        // trial sites, FP64
  uint4 pfix[kMaxPts];           // same, fixed point
};
constexpr size_t kWaveArea = sizeof(WaveLds);

struct WidomArgs {
  BoxConst bc;
  const double* fw;        // SoA x,y,z,q [4][n_fw], Morton order of the screen cells
  const int* fw_type;      // same order
  const uint4* fw_fixT;    // fixed-point x,y,z + type (scan without a cell grid)
  int n_fw;
  int use_cells;
  int cell_n[3];
  int cell_cap;
  const uint4* ctab;       // [cells * cell_cap] fixed-point x,y,z + type, ~0u = empty
  const double* g;         // guests SoA x,y,z,q [4][guest_cap]
  const int* g_type;
  const int* g_molid;
  const uint4* g_fix;
  int guest_cap;
  GuestView gv;
  const double* ff64;      // {4 eps, sigma^2, shift, eps} per type pair
  int comp, new_molid;
  int M, chain;
  int site_type[kMaxMol];
  double site_q[kMaxMol];
  double tvec[kMaxMol][3];    // template vectors pos[1+a] - pos[0] of the first molecule slot
  int K, O;
  double beta;
  int bias;
  double excl, tail_de;
  uint64_t seed;
  uint32_t stream;
  int64_t first_index;
  int n_insert;
  int charged;
  int n_kact;
  const int4* kact;         // active k-vectors: kx, ky, kz, index into the S(k) arrays
  const double* ktemp_act;  // their prefactor * exp(-k^2/4a^2)/k^2 (x2 for kx > 0)
  const double2* sk_ads;
  const double2* sk_fw;
  int kstride;              // max kmax + 1
  const double* coul;       // real-space Coulomb table (fast form)
  int coul_n;               // doubles
  // LDS layout (bytes from the start of the dynamic area)
  int lds_thr, lds_coul, lds_tab, lds_wave, wave_bytes;
  graspa_widom_result* res;   // may be null
  double* rec_w;              // [n_insert] W
  double* rec_e;              // [n_insert] dE
  int* rec_f;                 // [n_insert] bit 0 first bead placed, bit 1 complete
  unsigned* ctr;              // work counter
  unsigned long long* stat;   // screened pairs, evaluated pairs
};

// Rotate_Quaternions (mc_utilities.h:423-457)
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

__device__ inline double2 cmul2(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
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

// Tests of trial site P (molecule site `site`) against one candidate atom per
// lane; true on a lane whose atom certainly overlaps the site.
template <bool CUBIC>
__device__ inline bool screen_hit(const BoxConst& bc, const float* __restrict__ thr, const int T, const int site,
                                  const uint4 at, const bool valid, const uint4 P)
{
  const float th = valid ? thr[site * T + (int)at.w] : -1.0f;
  const float rr = r2_fixed32<CUBIC>(at.x, at.y, at.z, P.x, P.y, P.z, bc.h32);
  return rr < th;
}

// Overlap screen of n_trials trials x n_sites sites (trial site p = a *
// n_trials + t in W.pfix, molecule site site0 + a).  With the cell grid, a
// trial site is tested against the atoms of the 27 cells around its own cell,
// which hold every atom that can overlap it; otherwise against all atoms.
// Guests are always scanned.  Returns the mask of trials with a certain
// overlap.
template <bool CUBIC>
__device__ inline uint64_t screen(const WidomArgs& A, const uint4* __restrict__ tab, const float* __restrict__ thr,
                                  const WaveLds& W, const int n_trials, const int n_sites, const int site0,
                                  unsigned long long& stat)
{
  const BoxConst& bc = A.bc;
  const int T = bc.ntypes;
  const int lane = threadIdx.x & 63;
  const uint64_t all = (n_trials >= 64) ? ~0ull : ((1ull << n_trials) - 1ull);
  uint64_t alive = all;
  if (A.use_cells) {
    // this lane's (neighbour cell, slot) in each of the two lane groups
    const int C = A.cell_cap, nslot = 27 * C;
    int off[2][3], slot[2];
    bool lv[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int l = g * 64 + lane;
      lv[g] = l < nslot;
      const int nb = lv[g] ? l / C : 0;
      slot[g] = lv[g] ? l - nb * C : 0;
      off[g][0] = nb % 3 - 1;
      off[g][1] = (nb / 3) % 3 - 1;
      off[g][2] = nb / 9 - 1;
    }
    const int ngroups = nslot > 64 ? 2 : 1;
    const int n0 = A.cell_n[0], n1 = A.cell_n[1], n2 = A.cell_n[2];
    // four trial sites per step, so that their LDS gathers are in flight together
    constexpr int kStep = 4;
    for (int a = 0; a < n_sites; ++a) {
      for (int t0 = 0; t0 < n_trials; t0 += kStep) {
        if (!((alive >> t0) & 0xFull)) continue;
        uint4 P[kStep];
        int cell[kStep][3];
#pragma unroll
        for (int g = 0; g < kStep; ++g) {
          P[g] = W.pfix[a * n_trials + min(t0 + g, n_trials - 1)];
          cell[g][0] = (int)(((uint64_t)P[g].x * (uint64_t)n0) >> 32);
          cell[g][1] = (int)(((uint64_t)P[g].y * (uint64_t)n1) >> 32);
          cell[g][2] = (int)(((uint64_t)P[g].z * (uint64_t)n2) >> 32);
        }
        for (int gg = 0; gg < ngroups; ++gg) {
          uint4 at[kStep];
#pragma unroll
          for (int g = 0; g < kStep; ++g) {
            int x = cell[g][0] + off[gg][0], y = cell[g][1] + off[gg][1], z = cell[g][2] + off[gg][2];
            x = x < 0 ? x + n0 : (x >= n0 ? x - n0 : x);
            y = y < 0 ? y + n1 : (y >= n1 ? y - n1 : y);
            z = z < 0 ? z + n2 : (z >= n2 ? z - n2 : z);
            at[g] = tab[(x + n0 * (y + n1 * z)) * C + slot[gg]];
          }
#pragma unroll
          for (int g = 0; g < kStep; ++g) {
            const int t = t0 + g;
            const bool act = t < n_trials && ((alive >> t) & 1ull);
            const bool valid = lv[gg] && at[g].w != ~0u;
            if (act) stat += (unsigned long long)__popcll(__ballot(valid));
            if (__any(act && screen_hit<CUBIC>(bc, thr, T, site0 + a, at[g], valid, P[g]))) alive &= ~(1ull << t);
          }
        }
      }
    }
  } else {
    for (int c0 = 0; c0 < A.n_fw && alive; c0 += 64) {
      const int j = c0 + lane;
      const bool valid = j < A.n_fw;
      const uint4 f = valid ? A.fw_fixT[j] : make_uint4(0u, 0u, 0u, 0u);
      const unsigned nval = (unsigned)min(64, A.n_fw - c0);
      for (int a = 0; a < n_sites; ++a) {
        for (int t = 0; t < n_trials; ++t) {
          if (!((alive >> t) & 1ull)) continue;
          stat += nval;
          if (__any(screen_hit<CUBIC>(bc, thr, T, site0 + a, f, valid, W.pfix[a * n_trials + t])))
            alive &= ~(1ull << t);
        }
      }
    }
  }
  // guest molecules
  for (int c = 1; c < A.gv.n_comp && alive; ++c) {
    const int n = A.gv.seg[c].n_atoms, base = A.gv.seg[c].base;
    for (int c0 = 0; c0 < n && alive; c0 += 64) {
      const int jj = c0 + lane;
      bool valid = jj < n;
      uint4 f = make_uint4(0u, 0u, 0u, 0u);
      if (valid) {
        const int j = base + jj;
        if (c == A.comp && A.g_molid[j] == A.new_molid) valid = false;
        f = A.g_fix[j];
        f.w = (uint32_t)A.g_type[j];
      }
      const unsigned nval = (unsigned)min(64, n - c0);
      for (int a = 0; a < n_sites; ++a) {
        for (int t = 0; t < n_trials; ++t) {
          if (!((alive >> t) & 1ull)) continue;
          stat += nval;
          if (__any(screen_hit<CUBIC>(bc, thr, T, site0 + a, f, valid, W.pfix[a * n_trials + t])))
            alive &= ~(1ull << t);
        }
      }
    }
  }
  return all & ~alive;
}

// erfc(alpha r)/r from the per-box table (coulomb_table.cpp): the interval is
// read from the exponent and top mantissa bits of r^2 as FP32, the value is a
// degree-5 polynomial in r^2 - s_i.  Clamped indices only occur for r^2 below
// 2^-7, i.e. pairs the overlap rule already rejects.
__device__ inline double coul_table(const double2* __restrict__ ct, const int nint, const double rr)
{
  constexpr int kShift = 23 - kCoulTabMBits;
  const uint32_t bits = __float_as_uint((float)rr);
  int idx = (int)(bits >> kShift) - ((127 + kCoulTabELo) << kCoulTabMBits);
  idx = min(max(idx, 0), nint - 1);
  const double si = (double)__uint_as_float(bits & ~((1u << kShift) - 1u));
  const double d = rr - si;
  const double2* c = ct + 3 * idx;
  const double2 c01 = c[0], c23 = c[1], c45 = c[2];
  double g = fma(c45.y, d, c45.x);
  g = fma(g, d, c23.y);
  g = fma(g, d, c23.x);
  g = fma(g, d, c01.y);
  return fma(g, d, c01.x);
}

// Full energies of trial site p (molecule site `site`) against the framework
// and the guests: wave sums of LJ and real-space Coulomb, and the exact
// overlap flag (VDW_Coulomb.cu:1183-1352 for one trial site).  The framework
// is read in the Morton order of the screen cells, so the cutoff branches of
// a wave's 64 atoms mostly agree.
template <bool MIXED, bool CUBIC>
__device__ inline void eval_site(const WidomArgs& A, const double4* __restrict__ ffl, const double2* __restrict__ ct,
                                 const WaveLds& W, const int p, const int site, double& sv, double& sr, bool& ov,
                                 unsigned long long& stat)
{
  const BoxConst& bc = A.bc;
  const int nint = A.coul_n / kCoulTabStride;
  const int T = bc.ntypes;
  const int lane = threadIdx.x & 63;
  const double sq = A.site_q[site];
  const bool coul = bc.use_charges && sq != 0.0;
  const double px = W.pos[p][0], py = W.pos[p][1], pz = W.pos[p][2];
  const double4* fr = ffl + site * T;
  const double* fx = A.fw;
  const double* fq = A.fw + 3 * (size_t)A.n_fw;
  double v = 0.0, r = 0.0;
  bool o = false;
  auto pair = [&](double ax, double ay, double az, double aq, int ty, double& v_, double& r_, bool& o_) {
    double dx = ax - px, dy = ay - py, dz = az - pz;
    if (MIXED)
      pbc_rint<CUBIC>(dx, dy, dz, bc.cell, bc.inv);
    else
      pbc_fp64(dx, dy, dz, bc.cell, bc.inv, CUBIC ? 1 : 0);
    const double rr = dx * dx + dy * dy + dz * dz;
    if (rr < bc.cut_vdw2) {
      const double4 ff = fr[ty];
      const double e = MIXED ? lj_mixed(rr, ff.x, ff.y, ff.z) : lj_fp64(rr, ff.x, ff.y, ff.z);
      o_ = o_ || (e > bc.overlap) || (rr < 0.01);
      v_ += e;
    }
    if (coul && rr < bc.cut_coul2) {
      if (MIXED) {
        r_ += (bc.prefactor * aq * sq) * coul_table(ct, nint, rr);
      } else {
        const double rd = sqrt(rr);
        r_ += bc.prefactor * 1.0 * (aq * sq * erfc(bc.alpha * rd)) / rd;
      }
    }
  };
  // Two atoms per lane per iteration (independent chains), and the next pair
  // of atoms is loaded one iteration ahead: the framework stream does not fit
  // the L1 cache, so its loads are served from L2.
  const int n = A.n_fw;
  double nx0 = 0.0, ny0 = 0.0, nz0 = 0.0, nq0 = 0.0, nx1 = 0.0, ny1 = 0.0, nz1 = 0.0, nq1 = 0.0;
  int nt0 = 0, nt1 = 0;
  if (lane < n) { nx0 = fx[lane]; ny0 = fx[n + lane]; nz0 = fx[2 * n + lane]; nq0 = fq[lane]; nt0 = A.fw_type[lane]; }
  if (lane + 64 < n) {
    const int k = lane + 64;
    nx1 = fx[k]; ny1 = fx[n + k]; nz1 = fx[2 * n + k]; nq1 = fq[k]; nt1 = A.fw_type[k];
  }
  double v1 = 0.0, r1 = 0.0;
  bool o1 = false;
  for (int j = lane; j < n; j += 128) {
    const double ax0 = nx0, ay0 = ny0, az0 = nz0, aq0 = nq0, ax1 = nx1, ay1 = ny1, az1 = nz1, aq1 = nq1;
    const int ty0 = nt0, ty1 = nt1;
    const bool has1 = j + 64 < n;
    const int jn = j + 128;
    if (jn < n) { nx0 = fx[jn]; ny0 = fx[n + jn]; nz0 = fx[2 * n + jn]; nq0 = fq[jn]; nt0 = A.fw_type[jn]; }
    if (jn + 64 < n) {
      const int k = jn + 64;
      nx1 = fx[k]; ny1 = fx[n + k]; nz1 = fx[2 * n + k]; nq1 = fq[k]; nt1 = A.fw_type[k];
    }
    pair(ax0, ay0, az0, aq0, ty0, v, r, o);
    if (has1) pair(ax1, ay1, az1, aq1, ty1, v1, r1, o1);
  }
  v += v1;
  r += r1;
  o = o || o1;
  unsigned long long npair = (unsigned long long)A.n_fw;
  for (int c = 1; c < A.gv.n_comp; ++c) {
    const int n = A.gv.seg[c].n_atoms, base = A.gv.seg[c].base;
    npair += (unsigned long long)n;
    for (int jj = lane; jj < n; jj += 64) {
      const int j = base + jj;
      if (c == A.comp && A.g_molid[j] == A.new_molid) continue;
      const int ta = A.g_type[j];
      const double gq = A.g[3 * A.guest_cap + j];
      double dx = A.g[j] - px, dy = A.g[A.guest_cap + j] - py, dz = A.g[2 * A.guest_cap + j] - pz;
      if (MIXED)
        pbc_rint<CUBIC>(dx, dy, dz, bc.cell, bc.inv);
      else
        pbc_fp64(dx, dy, dz, bc.cell, bc.inv, CUBIC ? 1 : 0);
      const double rr = dx * dx + dy * dy + dz * dz;
      if (rr < bc.cut_vdw2) {
        const double4 ff = fr[ta];
        const double e = MIXED ? lj_mixed(rr, ff.x, ff.y, ff.z) : lj_fp64(rr, ff.x, ff.y, ff.z);
        o = o || (e > bc.overlap) || (rr < 0.01);
        v += e;
      }
      if (coul && rr < bc.cut_coul2) {
        if (MIXED) {
          r += (bc.prefactor * gq * sq) * coul_table(ct, nint, rr);
        } else {
          const double rd = sqrt(rr);
          r += bc.prefactor * 1.0 * (gq * sq * erfc(bc.alpha * rd)) / rd;
        }
      }
    }
  }
  stat += npair;
  sv = wave_sum(v);
  sr = coul ? wave_sum(r) : 0.0;
  ov = __any(o) != 0;
}

// SelectTrialPosition / CBMC_FirstBead_Finish / Widom_Move_Chain_PARTIAL
// (mc_widom.h:14-39, 305-383, 574-614): lane t < n holds trial t's LJ (ev) and
// real-space (er) sums; ovm marks overlapping trials.  Survivors, log-sum-exp
// shift, cumulative pick with the uniform u in trial order, Rosenbluth sum
// averaged over n, and the LJ-only bias correction.  Returns the selected
// trial (or -1) and W, both uniform.
__device__ inline void wave_select(double ev, double er, uint64_t ovm, int n, double u, double beta, int bias,
                                   int& sel_out, double& W_out)
{
  const int lane = threadIdx.x & 63;
  const bool alive = lane < n && !((ovm >> lane) & 1ull);
  double lw = -INFINITY;
  if (alive) {
    double tot = ev;
    if (bias) tot += er;
    lw = -beta * tot;
  }
  const double mx = wave_max(lw);
  const double shf = alive ? exp(lw - mx) : -1.0;   // shifted factors; -1 marks an overlap
  const double ew = alive ? exp(lw) : 0.0;
  int nk = 0;
  double ssum = 0.0, rb = 0.0;
  for (int t = 0; t < n; ++t) {
    const double s = readlane_d(shf, t);
    if (s >= 0.0) {
      ssum += s;
      rb += readlane_d(ew, t);
      ++nk;
    }
  }
  int sel = -1;
  if (nk > 0) {
    const double ws = u * ssum;
    int t = 0;
    while (readlane_d(shf, t) < 0.0) ++t;
    double cumw = readlane_d(shf, t);
    sel = t;
    int left = nk - 1;
    while (cumw < ws && left > 0) {
      ++t;
      while (readlane_d(shf, t) < 0.0) ++t;
      cumw += readlane_d(shf, t);
      sel = t;
      --left;
    }
  }
  double W = 0.0;
  if (nk > 0 && !(rb < 1e-150)) {
    W = rb / (double)n;
    if (!bias) W *= exp(-beta * readlane_d(er, sel));
  } else {
    sel = -1;
  }
  sel_out = sel;
  W_out = W;
}

template <bool MIXED, bool TAB_LDS, bool CUBIC>
__global__ __launch_bounds__(kMaxWaves * 64) void widom_kernel(const WidomArgs A)
{
  extern __shared__ __align__(16) char dyn[];
  const BoxConst& bc = A.bc;
  const int T = bc.ntypes;
  const int M = A.M, chain = A.chain, K = A.K, O = A.O;
  // per-(site, atom type) force-field rows and screen thresholds
  double4* ffl = reinterpret_cast<double4*>(dyn);
  float* thr = reinterpret_cast<float*>(dyn + A.lds_thr);
  for (int i = threadIdx.x; i < M * T; i += blockDim.x) {
    const int a = i / T, ty = i % T;
    const double* f = A.ff64 + 4 * (ty * T + A.site_type[a]);
    ffl[i] = make_double4(f[0], f[1], f[2], 0.0);
    thr[i] = screen_threshold(f, bc.overlap, bc.cut_vdw2);
  }
  // real-space Coulomb table of the fast form
  double2* ctl = reinterpret_cast<double2*>(dyn + A.lds_coul);
  if (MIXED)
    for (int i = threadIdx.x; i < A.coul_n / 2; i += blockDim.x) ctl[i] = reinterpret_cast<const double2*>(A.coul)[i];
  // the screen's cell table, in LDS when it fits
  const uint4* tab;
  if (TAB_LDS) {
    uint4* s = reinterpret_cast<uint4*>(dyn + A.lds_tab);
    const int n = A.cell_n[0] * A.cell_n[1] * A.cell_n[2] * A.cell_cap;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = A.ctab[i];
    tab = s;
  } else {
    tab = A.ctab;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  WaveLds& W = *reinterpret_cast<WaveLds*>(dyn + A.lds_wave + (size_t)wv * A.wave_bytes);
  double2* eik = reinterpret_cast<double2*>(&W);   // Ewald phase only
  double (*mol)[3] = reinterpret_cast<double (*)[3]>(dyn + A.lds_wave + (size_t)wv * A.wave_bytes + A.wave_bytes -
                                                     kMaxMol * 3 * sizeof(double));
  unsigned long long st_screen = 0, st_full = 0;
  const int nd = 3 * K + 1 + (chain > 0 ? 3 * O + 1 : 0);

  unsigned next = 0;
  if (lane == 0) next = atomicAdd(A.ctr, 1u);
  for (;;) {
    const int ins = __builtin_amdgcn_readfirstlane((int)next);
    if (ins >= A.n_insert) break;
    if (lane == 0) next = atomicAdd(A.ctr, 1u);
    const uint64_t idx = (uint64_t)(A.first_index + ins);

    // 1. draws: block b holds draws 2b and 2b+1 (philox_uniform's layout)
    for (int b = lane; 2 * b < nd; b += 64) {
      uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)b, A.stream};
      philox4x32_10(c, (uint32_t)A.seed, (uint32_t)(A.seed >> 32));
      const uint64_t x0 = (((uint64_t)c[0] << 32) | c[1]) >> 11;
      W.ud[2 * b] = ((double)x0 + 0.5) * (1.0 / 9007199254740992.0);
      if (2 * b + 1 < nd) {
        const uint64_t x1 = (((uint64_t)c[2] << 32) | c[3]) >> 11;
        W.ud[2 * b + 1] = ((double)x1 + 0.5) * (1.0 / 9007199254740992.0);
      }
    }
    wave_sync();
    // 2. first-bead trial positions: diag(cell) * U^3 (get_random_trial_position, mc_widom.h:122-213)
    if (lane < K) {
      double p[3];
      p[0] = bc.cell[0] * W.ud[3 * lane];
      p[1] = bc.cell[4] * W.ud[3 * lane + 1];
      p[2] = bc.cell[8] * W.ud[3 * lane + 2];
      W.pos[lane][0] = p[0]; W.pos[lane][1] = p[1]; W.pos[lane][2] = p[2];
      W.pfix[lane] = to_fixed4(bc, p);
    }
    wave_sync();
    // 3.-5. first bead
    uint64_t ovm = screen<CUBIC>(A, tab, thr, W, K, 1, 0, st_screen);
    double ev = 0.0, er = 0.0;
    for (int t = 0; t < K; ++t) {
      if ((ovm >> t) & 1ull) continue;
      double sv, sr;
      bool o;
      eval_site<MIXED, CUBIC>(A, ffl, ctl, W, t, 0, sv, sr, o, st_full);
      if (o) ovm |= 1ull << t;
      if (lane == t) { ev = sv; er = sr; }
    }
    int sel;
    double Wfb;
    wave_select(ev, er, ovm, K, W.ud[3 * K], A.beta, A.bias, sel, Wfb);
    const bool fb_ok = sel >= 0 && Wfb > 1e-150;
    bool ok = fb_ok;
    double Wtot = Wfb, e_v = 0.0, e_r = 0.0;
    int orient = 0;
    if (fb_ok) {
      e_v = readlane_d(ev, sel);
      e_r = readlane_d(er, sel);
      if (lane < 3) mol[0][lane] = W.pos[sel][lane];
    }
    wave_sync();
    // 6. orientations of the template about the selected first bead
    //    (get_random_trial_orientation, mc_widom.h:215-303); trial site of
    //    orientation o and chain atom a at index a * O + o
    if (fb_ok && chain > 0) {
      if (lane < O) {
        const double u = W.ud[3 * K + 1 + 3 * lane];
        const double v = W.ud[3 * K + 2 + 3 * lane];
        const double w = W.ud[3 * K + 3 + 3 * lane];
        for (int a = 0; a < chain; ++a) {
          double vec[3] = {A.tvec[a][0], A.tvec[a][1], A.tvec[a][2]};
          quat_rotate(vec, u, v, w);
          const int s = a * O + lane;
          double p[3] = {mol[0][0] + vec[0], mol[0][1] + vec[1], mol[0][2] + vec[2]};
          W.pos[s][0] = p[0]; W.pos[s][1] = p[1]; W.pos[s][2] = p[2];
          W.pfix[s] = to_fixed4(bc, p);
        }
      }
      wave_sync();
      uint64_t ovo = screen<CUBIC>(A, tab, thr, W, O, chain, 1, st_screen);
      double cv = 0.0, cr = 0.0;
      for (int o = 0; o < O; ++o) {
        if ((ovo >> o) & 1ull) continue;
        double tv = 0.0, tr = 0.0;
        bool dead = false;
        for (int a = 0; a < chain && !dead; ++a) {
          double sv, sr;
          bool ob;
          eval_site<MIXED, CUBIC>(A, ffl, ctl, W, a * O + o, 1 + a, sv, sr, ob, st_full);
          tv += sv;
          tr += sr;
          dead = ob;
        }
        if (dead) ovo |= 1ull << o;
        if (lane == o) { cv = tv; cr = tr; }
      }
      int so;
      double Wch;
      wave_select(cv, cr, ovo, O, W.ud[3 * K + 3 * O + 1], A.beta, A.bias, so, Wch);
      if (so < 0 || !(Wfb * Wch > 1e-150)) {
        ok = false;
      } else {
        Wtot = Wfb * Wch;
        orient = so;
        e_v += readlane_d(cv, so);
        e_r += readlane_d(cr, so);
        for (int i = lane; i < 3 * chain; i += 64) {
          const int a = i / 3, d = i % 3;
          mol[1 + a][d] = W.pos[a * O + so][d];
        }
      }
      wave_sync();
    }
    // 7. Fourier-space energy of the selected molecule, FP64 (Fourier_Ewald_Diff,
    //    Ewald_Energy_Functions.h:280-397, with Oldsize = 0)
    double ew = 0.0;
    if (ok && A.charged) {
      const int KS = A.kstride;
      for (int t = lane; t < M * 3; t += 64) {
        const int site = t / 3, axis = t % 3;
        double s = bc.inv[axis] * mol[site][0] + bc.inv[3 + axis] * mol[site][1] + bc.inv[6 + axis] * mol[site][2];
        s *= 2 * 3.14159265358979323846;
        double sn, cs;
        sincos(s, &sn, &cs);
        const double2 e1 = make_double2(cs, sn);
        double2* e = eik + (size_t)t * KS;
        e[0] = make_double2(1.0, 0.0);
        if (bc.kmax[axis] >= 1) e[1] = e1;
        double2 cur = e1;
        for (int k = 2; k <= bc.kmax[axis]; ++k) {
          cur = cmul2(cur, e1);
          e[k] = cur;
        }
      }
      wave_sync();
      double es = 0.0, ec = 0.0;
      for (int kk = lane; kk < A.n_kact; kk += 64) {
        const int4 kv = A.kact[kk];
        const double t = A.ktemp_act[kk];
        const int aky = abs(kv.y), akz = abs(kv.z);
        double2 cn = make_double2(0.0, 0.0);
        for (int i = 0; i < M; ++i) {
          const double2* e = eik + (size_t)(3 * i) * KS;
          double2 ey = e[KS + aky];
          if (kv.y < 0) ey.y = -ey.y;
          const double2 exy = cmul2(e[kv.x], ey);
          double2 ez = e[2 * KS + akz];
          if (kv.z < 0) ez.y = -ez.y;
          const double2 pr = cmul2(exy, ez);
          cn.x += A.site_q[i] * pr.x;
          cn.y += A.site_q[i] * pr.y;
        }
        const double2 ov = A.sk_ads[kv.w];
        const double2 nv = make_double2(ov.x + cn.x, ov.y + cn.y);
        es += t * (nv.x * nv.x + nv.y * nv.y) - t * (ov.x * ov.x + ov.y * ov.y);
        const double2 cvk = A.sk_fw[kv.w];
        ec += t * (cvk.x * cn.x + cvk.y * cn.y);
      }
      es = wave_sum(es);
      ec = wave_sum(ec);
      ew = (es - A.excl) + 2.0 * ec;
      wave_sync();
    }
    // 8. record
    if (lane == 0) {
      double Wf = 0.0, etot = 0.0;
      if (ok) {
        Wf = Wtot * exp(-A.beta * ew) * exp(-A.beta * A.tail_de);
        etot = e_v + e_r + ew + A.tail_de;
      }
      A.rec_w[ins] = Wf;
      A.rec_e[ins] = etot;
      A.rec_f[ins] = (fb_ok ? 1 : 0) | (ok ? 2 : 0);
      if (A.res) {
        graspa_widom_result r;
        r.rosen = Wf;
        r.energy = etot;
        r.vdw = ok ? e_v : 0.0;
        r.real = ok ? e_r : 0.0;
        r.ewald = ok ? ew : 0.0;
        r.first_bead = ok ? sel : -1;
        r.orientation = ok ? orient : -1;
        r.ok = ok ? 1 : 0;
        r._pad = 0;
        A.res[ins] = r;
      }
    }
  }
  if (lane == 0) {
    atomicAdd(&A.stat[0], st_screen);
    atomicAdd(&A.stat[1], st_full);
  }
}

__global__ void widom_reset_kernel(unsigned* ctr, unsigned long long* stat)
{
  if (threadIdx.x == 0) {
    *ctr = 0u;
    stat[0] = 0ull;
    stat[1] = 0ull;
  }
}

// Sums of the per-insertion records over fixed index ranges, in a fixed order
// (deterministic): block b covers records [b*2048, (b+1)*2048).
__global__ __launch_bounds__(kRedBlock) void widom_reduce_kernel(const double* rw, const double* re, const int* rf,
                                                                 int n, double* part)
{
  __shared__ double s[kSums][kRedBlock];
  double acc[kSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int base = blockIdx.x * kRedBlock * kRedPer;
  for (int k = 0; k < kRedPer; ++k) {
    const int i = base + k * kRedBlock + threadIdx.x;
    if (i < n) {
      const double w = rw[i];
      const int f = rf[i];
      acc[0] += w;
      acc[1] += w * w;
      acc[2] += w * re[i];
      acc[3] += 1.0;
      acc[4] += (f & 1) ? 1.0 : 0.0;
      acc[5] += (f & 2) ? 1.0 : 0.0;
    }
  }
  for (int q = 0; q < kSums; ++q) s[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int h = kRedBlock / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int q = 0; q < kSums; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < kSums) part[kSums * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(kRedBlock) void widom_final_kernel(const double* part, int nblk, double* out)
{
  __shared__ double s[kSums][kRedBlock];
  double acc[kSums] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += kRedBlock)
    for (int q = 0; q < kSums; ++q) acc[q] += part[kSums * b + q];
  for (int q = 0; q < kSums; ++q) s[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int h = kRedBlock / 2; h >= 1; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int q = 0; q < kSums; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < kSums) out[threadIdx.x] = s[threadIdx.x][0];
}

// Waves per workgroup and workgroups per CU for one instantiation: the most
// resident waves, ties to the larger workgroup (fewer cell-table copies).
template <bool MIXED, bool TAB_LDS, bool CUBIC>
static int launch_one(graspa_box* b, const WidomArgs& A0, size_t fixed_lds, size_t wave_bytes)
{
  auto kern = widom_kernel<MIXED, TAB_LDS, CUBIC>;
  int best_nw = 0, best_blocks = 0, best_waves = 0;
  size_t best_lds = 0;
  for (int nw : {4, kMaxWaves}) {
    const size_t lds = fixed_lds + (size_t)nw * wave_bytes;
    if (lds > (size_t)b->max_lds) continue;
    if (lds > 65536) {
      (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      (void)hipGetLastError();
    }
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kern, 64 * nw, lds) != hipSuccess || blocks < 1) {
      (void)hipGetLastError();
      continue;
    }
    if (blocks * nw >= best_waves) {
      best_waves = blocks * nw;
      best_nw = nw;
      best_blocks = blocks;
      best_lds = lds;
    }
  }
  if (best_nw == 0) return fail(GRASPA_ERR_UNSUPPORTED, "widom: no launch shape fits the LDS");
  WidomArgs A = A0;
  const long long waves_needed = (A.n_insert + 0LL);
  long long grid = (long long)best_blocks * b->n_cu;
  grid = std::max(1LL, std::min(grid, (waves_needed + best_nw - 1) / best_nw));
  timing_begin(b, "widom");
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * best_nw), best_lds, b->stream, A);
  timing_end(b, "widom");
  if (hip_check(hipGetLastError(), "widom_kernel launch")) return GRASPA_ERR_HIP;
  return GRASPA_OK;
}

template <bool MIXED>
static int launch_mixed(graspa_box* b, const WidomArgs& A, bool tab_lds, size_t fixed_lds, size_t wave_bytes)
{
  if (tab_lds)
    return A.bc.cubic ? launch_one<MIXED, true, true>(b, A, fixed_lds, wave_bytes)
                      : launch_one<MIXED, true, false>(b, A, fixed_lds, wave_bytes);
  return A.bc.cubic ? launch_one<MIXED, false, true>(b, A, fixed_lds, wave_bytes)
                    : launch_one<MIXED, false, false>(b, A, fixed_lds, wave_bytes);
}

static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

int launch_widom(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                 uint64_t seed, uint32_t stream, int64_t first_index, int n_insert, graspa_widom_result* d_res,
                 graspa_widom_summary* summary)
{
  if (comp < 1 || comp >= (int)b->comps.size()) return fail(GRASPA_ERR_ARG, "widom: bad component");
  const Component& c = b->comps[comp];
  if (K < 1 || K > kMaxTrials || O < 1 || O > kMaxTrials) return fail(GRASPA_ERR_ARG, "widom: 1 <= K,O <= 32");
  if (c.molsize < 1 || c.molsize > kMaxMol) return fail(GRASPA_ERR_UNSUPPORTED, "widom: molecule larger than 8 sites");
  if (O * (c.molsize - 1) > kMaxPts) return fail(GRASPA_ERR_UNSUPPORTED, "widom: orientations x chain sites above 64");
  if (b->bc.ntypes > kMaxTypes) return fail(GRASPA_ERR_UNSUPPORTED, "widom: too many pseudo-atom types");
  if (n_insert <= 0) {
    if (summary) *summary = graspa_widom_summary{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    return GRASPA_OK;
  }
  if (ensure_widom_records(b, n_insert)) return GRASPA_ERR_HIP;
  WidomArgs A{};
  A.bc = b->bc;
  A.fw = b->d_fws; A.fw_type = b->d_fws_type; A.fw_fixT = b->d_fw_fixT; A.n_fw = b->n_fw;
  A.use_cells = b->use_cells;
  for (int d = 0; d < 3; ++d) A.cell_n[d] = b->cell_n[d];
  A.cell_cap = b->cell_cap;
  A.ctab = b->d_ctab;
  A.g = b->d_g; A.g_type = b->d_g_type; A.g_molid = b->d_g_molid; A.g_fix = b->d_g_fix; A.guest_cap = b->guest_cap;
  guest_view(b, &A.gv);
  A.ff64 = b->d_ff64;
  A.comp = comp;
  A.new_molid = c.n_mol;
  A.M = c.molsize;
  A.chain = c.molsize - 1;
  for (int a = 0; a < c.molsize; ++a) {
    A.site_type[a] = c.type[a];
    A.site_q[a] = c.q[a];
    if (a >= 1)
      for (int d = 0; d < 3; ++d) A.tvec[a - 1][d] = c.pos[3 * a + d] - c.pos[d];
  }
  double qabs = 0.0;
  for (int a = 0; a < c.molsize; ++a) qabs += std::fabs(c.q[a]);
  A.charged = (b->bc.use_charges && qabs > 1e-6) ? 1 : 0;
  A.K = K; A.O = O; A.beta = beta; A.bias = bias; A.excl = excl; A.tail_de = tail_de;
  A.seed = seed; A.stream = stream; A.first_index = first_index; A.n_insert = n_insert;
  A.n_kact = b->n_active;
  A.kact = b->d_kact;
  A.ktemp_act = b->d_ktemp_act;
  A.sk_ads = b->d_sk_ads; A.sk_fw = b->d_sk_fw;
  A.kstride = 1 + std::max(b->bc.kmax[0], std::max(b->bc.kmax[1], b->bc.kmax[2]));
  A.res = d_res;
  A.rec_w = b->d_wrec_w; A.rec_e = b->d_wrec_e; A.rec_f = b->d_wrec_f;
  A.ctr = reinterpret_cast<unsigned*>(b->d_wacc + 12);
  A.stat = reinterpret_cast<unsigned long long*>(b->d_wacc + 8);
  A.coul = b->d_coul;
  A.coul_n = b->coul_n;
  // LDS layout: FF rows | thresholds | Coulomb table (fast form) | cell table (optional) | per-wave areas
  const int T = A.bc.ntypes, M = A.M;
  size_t off = align16((size_t)M * T * sizeof(double4));
  A.lds_thr = (int)off;
  off = align16(off + (size_t)M * T * sizeof(float));
  A.lds_coul = (int)off;
  if (b->precision == GRASPA_MIXED) off = align16(off + (size_t)b->coul_n * sizeof(double));
  const size_t eik_bytes = (size_t)M * 3 * A.kstride * sizeof(double2);
  const size_t wave_bytes = align16(std::max(kWaveArea, eik_bytes)) + kMaxMol * 3 * sizeof(double);
  A.wave_bytes = (int)wave_bytes;
  const size_t tab_bytes =
      b->use_cells ? (size_t)b->cell_n[0] * b->cell_n[1] * b->cell_n[2] * b->cell_cap * sizeof(uint4) : 0;
  A.lds_tab = (int)off;
  const size_t base_lds = off;

  hipLaunchKernelGGL(widom_reset_kernel, dim3(1), dim3(64), 0, b->stream, A.ctr, A.stat);
  if (hip_check(hipGetLastError(), "widom_reset launch")) return GRASPA_ERR_HIP;
  // cell table in LDS when it fits beside four wave areas, else read through the caches
  int rc = GRASPA_ERR_UNSUPPORTED;
  for (int pass = 0; pass < 2 && rc == GRASPA_ERR_UNSUPPORTED; ++pass) {
    const bool tab_lds = pass == 0;
    if (tab_lds && (!b->use_cells || base_lds + align16(tab_bytes) + 4 * wave_bytes > (size_t)b->max_lds)) continue;
    const size_t fixed_lds = tab_lds ? base_lds + align16(tab_bytes) : base_lds;
    A.lds_wave = (int)fixed_lds;
    rc = (b->precision == GRASPA_MIXED) ? launch_mixed<true>(b, A, tab_lds, fixed_lds, wave_bytes)
                                        : launch_mixed<false>(b, A, tab_lds, fixed_lds, wave_bytes);
  }
  if (rc) return rc;
  const int nblk = (n_insert + kRedBlock * kRedPer - 1) / (kRedBlock * kRedPer);
  hipLaunchKernelGGL(widom_reduce_kernel, dim3(nblk), dim3(kRedBlock), 0, b->stream, A.rec_w, A.rec_e, A.rec_f,
                     n_insert, b->d_wpart);
  if (hip_check(hipGetLastError(), "widom_reduce launch")) return GRASPA_ERR_HIP;
  hipLaunchKernelGGL(widom_final_kernel, dim3(1), dim3(kRedBlock), 0, b->stream, b->d_wpart, nblk, b->d_wacc);
  if (hip_check(hipGetLastError(), "widom_final launch")) return GRASPA_ERR_HIP;
  if (summary) {
    double acc[12];
    if (hip_check(hipMemcpyAsync(acc, b->d_wacc, 12 * sizeof(double), hipMemcpyDeviceToHost, b->stream),
                  "widom summary copy"))
      return GRASPA_ERR_HIP;
    if (hip_check(hipStreamSynchronize(b->stream), "widom sync")) return GRASPA_ERR_HIP;
    unsigned long long st[2];
    std::memcpy(st, acc + 8, sizeof(st));
    summary->sum_w = acc[0];
    summary->sum_w2 = acc[1];
    summary->sum_we = acc[2];
    summary->n = acc[3];
    summary->n_first_bead = acc[4];
    summary->n_ok = acc[5];
    summary->screened_pairs = (double)st[0];
    summary->evaluated_pairs = (double)st[1];
  }
  return GRASPA_OK;
}

}  // namespace graspa
