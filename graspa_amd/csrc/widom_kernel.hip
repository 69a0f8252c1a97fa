// widom_kernel.hip -- batched Widom / CBMC test insertions in one persistent
// launch (the throughput path of the GCMC energy engine on gfx950).
//
// Each workgroup (4 x wave64) keeps its stripe of the framework in registers
// for the whole launch (atom j of the stripe lives in lane j mod 256), and walks
// insertions i = blockIdx.x, blockIdx.x + gridDim.x, ...  One insertion:
//   1. K first-bead trial positions from the insertion's Philox stream (LDS);
//   2. every lane evaluates its atoms against all K trials, accumulating per-trial
//      LJ and real-space sums (FP32 pair terms in the mixed form), then one
//      wave64 FP64 reduction per trial and a cross-wave sum in LDS;
//   3. log-sum-exp selection of the first bead (one uniform), W_fb;
//   4. O rigid-body orientations of the template (quaternions from 3 uniforms),
//      the same pair pass over O x (M-1) sites, selection, W_ch;
//   5. Fourier-space energy of the selected molecule in FP64 against the
//      resident structure factors (one lane per k-vector);
//   6. W = W_fb W_ch exp(-beta (dE_ewald + dE_tail)), record, block sums.
// The semantics are Insertion_Body (mc_swap_utilities.h:3-133) with
// Widom_Move_FirstBead_PARTIAL / Widom_Move_Chain_PARTIAL (mc_widom.h:305-614).
#include "device_math.hpp"
#include "engine.hpp"

namespace graspa {

constexpr int kMaxTrials = 32;   // K and O
constexpr int kMaxMol = 8;       // sites per molecule
constexpr int kMaxKStride = 16;  // kmax <= 15 along each axis
constexpr int kSums = 6;         // W, W^2, W dE, n, n first bead placed, n complete

struct WidomArgs {
  BoxConst bc;
  const double* fw;        // SoA x,y,z,q
  const int* fw_type;
  const uint4* fw_fix;
  int n_fw;
  const double* g;
  const int* g_type;
  const int* g_molid;
  const uint4* g_fix;
  int guest_cap;
  GuestView gv;
  const double* ff64;
  const float* ff32;
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
  const double* ktemp;
  const double2* sk_ads;
  const double2* sk_fw;
  graspa_widom_result* res;   // may be null
  double* wsum;               // [gridDim.x*kSums + kSums]
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

__device__ inline void to_fixed(const BoxConst& bc, const double p[3], uint32_t f[3])
{
  double s[3];
  s[0] = bc.inv[0] * p[0] + bc.inv[3] * p[1] + bc.inv[6] * p[2];
  s[1] = bc.inv[1] * p[0] + bc.inv[4] * p[1] + bc.inv[7] * p[2];
  s[2] = bc.inv[2] * p[0] + bc.inv[5] * p[1] + bc.inv[8] * p[2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double fr = s[d] - floor(s[d]);
    f[d] = (uint32_t)(int64_t)(fr * 4294967296.0);
  }
}

__device__ inline double2 cmul2(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Shared workspace of one workgroup.
struct WidomShared {
  double pos[kMaxTrials * kMaxMol][3];      // trial sites (FP64)
  uint32_t fix[kMaxTrials * kMaxMol][3];    // same, fixed point
  double red[4][kMaxTrials][2];             // per-wave partial sums (LJ, real) per trial
  uint32_t ov[4];                           // per-wave overlap bits
  double tot[kMaxTrials][2];                // workgroup sums per trial
  double logw[kMaxTrials], ew[kMaxTrials];
  double W;
  int ok;
  double mol[kMaxMol][3];                   // selected molecule
  double2 eik[kMaxMol][3][kMaxKStride];
  double ered[4][2];
  double e_fb[2], e_ch[2];
  int fb, orient;
};

// Pair pass of this lane's framework atoms (registers) and of the guest atoms
// against n_trials trials of `chain` sites (sh.pos / sh.fix).  Each trial row
// is summed over the lane's atoms (FP32 terms in the mixed form), reduced over
// the wave in FP64 at once, and the wave sums added in LDS; overlap bits by
// wave vote.  Result: sh.tot[t] and sh.ov[0].
template <bool MIXED, int APT>
__device__ inline void pair_pass(const WidomArgs& A, WidomShared& sh, const int n_trials, const int chain,
                                 const double (&fx)[APT], const double (&fy)[APT], const double (&fz)[APT],
                                 const double (&fq)[APT], const uint32_t (&ux)[APT], const uint32_t (&uy)[APT],
                                 const uint32_t (&uz)[APT], const float (&uq)[APT], const int (&ft)[APT],
                                 const int n_mine, const float4* ffl32, const double* ffl64)
{
  const BoxConst& bc = A.bc;
  const int T = bc.ntypes;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane < n_trials) { sh.red[wave][lane][0] = 0.0; sh.red[wave][lane][1] = 0.0; }
  uint32_t ovbits = 0u;
  for (int a = 0; a < chain; ++a) {
    const int st = A.site_type[(chain == 1) ? 0 : 1 + a];
    const double sq = A.site_q[(chain == 1) ? 0 : 1 + a];
    // force-field rows and charge products of this lane's atoms for site type st
    double fa[APT], fs[APT], fh[APT], qq[APT];
    if (MIXED) {
#pragma unroll
      for (int j = 0; j < APT; ++j) {
        const double* f = ffl64 + 4 * (ft[j] * T + st);
        fa[j] = f[0]; fs[j] = f[1]; fh[j] = f[2];
        qq[j] = bc.prefactor * fq[j] * sq;
      }
    }
    for (int t = 0; t < n_trials; ++t) {
      const int s = t * chain + a;
      double sv = 0.0, sr = 0.0;
      bool ov = false;
      if (MIXED) {
        const uint32_t sx = sh.fix[s][0], sy = sh.fix[s][1], sz = sh.fix[s][2];
#pragma unroll
        for (int j = 0; j < APT; ++j) {
          if (j >= n_mine) continue;
          const double rr = r2_fixed64(ux[j], uy[j], uz[j], sx, sy, sz, bc.h64, bc.cubic);
          if (rr < bc.cut_vdw2) {
            const double ev = lj_mixed(rr, fa[j], fs[j], fh[j]);
            ov = ov || (ev > bc.overlap) || (rr < 0.01);
            sv += ev;
          }
          if (bc.use_charges && rr < bc.cut_coul2) sr += coulomb_mixed(rr, qq[j], bc.alpha);
        }
      } else {
        const double px = sh.pos[s][0], py = sh.pos[s][1], pz = sh.pos[s][2];
#pragma unroll
        for (int j = 0; j < APT; ++j) {
          if (j >= n_mine) continue;
          double dx = fx[j] - px, dy = fy[j] - py, dz = fz[j] - pz;
          pbc_fp64(dx, dy, dz, bc.cell, bc.inv, bc.cubic);
          const double rr = dx * dx + dy * dy + dz * dz;
          if (rr < bc.cut_vdw2) {
            const double* f = ffl64 + 4 * (ft[j] * T + st);
            const double e = lj_fp64(rr, f[0], f[1], f[2]);
            ov = ov || (e > bc.overlap) || (rr < 0.01);
            sv += e;
          }
          if (bc.use_charges && rr < bc.cut_coul2) {
            const double r = sqrt(rr);
            sr += bc.prefactor * 1.0 * (fq[j] * sq * erfc(bc.alpha * r)) / r;
          }
        }
      }
      // guest molecules (guest-guest), streamed from memory
      for (int c = 1; c < A.gv.n_comp; ++c) {
        const int n = A.gv.seg[c].n_atoms, base = A.gv.seg[c].base;
        for (int jj = threadIdx.x; jj < n; jj += blockDim.x) {
          const int j = base + jj;
          if (c == A.comp && A.g_molid[j] == A.new_molid) continue;
          const int ta = A.g_type[j];
          if (MIXED) {
            const uint4 g4 = A.g_fix[j];
            const double* f = ffl64 + 4 * (ta * T + st);
            const double rr = r2_fixed64(g4.x, g4.y, g4.z, sh.fix[s][0], sh.fix[s][1], sh.fix[s][2], bc.h64, bc.cubic);
            if (rr < bc.cut_vdw2) {
              const double ev = lj_mixed(rr, f[0], f[1], f[2]);
              ov = ov || (ev > bc.overlap) || (rr < 0.01);
              sv += ev;
            }
            if (bc.use_charges && rr < bc.cut_coul2)
              sr += coulomb_mixed(rr, bc.prefactor * A.g[3 * A.guest_cap + j] * sq, bc.alpha);
          } else {
            const double* f = ffl64 + 4 * (ta * T + st);
            double dx = A.g[j] - sh.pos[s][0], dy = A.g[A.guest_cap + j] - sh.pos[s][1];
            double dz = A.g[2 * A.guest_cap + j] - sh.pos[s][2];
            const double gq = A.g[3 * A.guest_cap + j];
            pbc_fp64(dx, dy, dz, bc.cell, bc.inv, bc.cubic);
            const double rr = dx * dx + dy * dy + dz * dz;
            if (rr < bc.cut_vdw2) {
              const double e = lj_fp64(rr, f[0], f[1], f[2]);
              ov = ov || (e > bc.overlap) || (rr < 0.01);
              sv += e;
            }
            if (bc.use_charges && rr < bc.cut_coul2) {
              const double r = sqrt(rr);
              sr += bc.prefactor * 1.0 * (gq * sq * erfc(bc.alpha * r)) / r;
            }
          }
        }
      }
      sv = wave_sum(sv);
      sr = bc.use_charges ? wave_sum(sr) : 0.0;
      const bool any = __any(ov);
      if (any) ovbits |= (1u << t);
      if (lane == 0) { sh.red[wave][t][0] += sv; sh.red[wave][t][1] += sr; }
    }
  }
  if (lane == 0) sh.ov[wave] = ovbits;
  __syncthreads();
  if (threadIdx.x < n_trials) {
    const int t = threadIdx.x;
    sh.tot[t][0] = sh.red[0][t][0] + sh.red[1][t][0] + sh.red[2][t][0] + sh.red[3][t][0];
    sh.tot[t][1] = sh.red[0][t][1] + sh.red[1][t][1] + sh.red[2][t][1] + sh.red[3][t][1];
  }
  if (threadIdx.x == 0) sh.ov[0] = sh.ov[0] | sh.ov[1] | sh.ov[2] | sh.ov[3];
  __syncthreads();
}

// CBMC_FirstBead_Finish / Widom_Move_Chain_PARTIAL selection (mc_widom.h:14-39,
// 305-383, 574-614): survivors, log-sum-exp shift, cumulative pick with the
// uniform u, Rosenbluth sum.  Called by the whole workgroup: wave 0 evaluates
// the exponentials in parallel, thread 0 sums in trial order as the reference
// does.  Leaves the selected trial (or -1) in *sel and the averaged weight in
// *W (shared variables).
__device__ inline void select_trial(const WidomArgs& A, WidomShared& sh, const int n_trials, const double u,
                                    double* W_out, int* sel_out)
{
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    const bool alive = lane < n_trials && !((sh.ov[0] >> lane) & 1u);
    double lw = -INFINITY;
    if (alive) {
      double tot = sh.tot[lane][0];
      if (A.bias) tot += sh.tot[lane][1];
      lw = -A.beta * tot;
    }
    double mx = lw;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) mx = fmax(mx, __shfl_xor(mx, m, 64));
    if (lane < n_trials) {
      sh.logw[lane] = alive ? exp(lw - mx) : -1.0;   // shifted factors; -1 marks an overlap
      sh.ew[lane] = alive ? exp(lw) : 0.0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int nk = 0, sel = -1;
    double ssum = 0.0, rb = 0.0;
    for (int t = 0; t < n_trials; ++t)
      if (sh.logw[t] >= 0.0) { ssum += sh.logw[t]; rb += sh.ew[t]; ++nk; }
    if (nk > 0) {
      const double ws = u * ssum;
      int t = 0;
      while (sh.logw[t] < 0.0) ++t;
      double cumw = sh.logw[t];
      sel = t;
      int left = nk - 1;
      while (cumw < ws && left > 0) {
        ++t;
        while (sh.logw[t] < 0.0) ++t;
        cumw += sh.logw[t];
        sel = t;
        --left;
      }
    }
    double W = 0.0;
    if (nk > 0 && !(rb < 1e-150)) {
      W = rb / (double)n_trials;
      if (!A.bias) W *= exp(-A.beta * sh.tot[sel][1]);
    } else {
      sel = -1;
    }
    *W_out = W;
    *sel_out = sel;
  }
  __syncthreads();
}

template <bool MIXED, int APT>
__global__ __launch_bounds__(kBlock, 1) void widom_kernel(const WidomArgs A)
{
  __shared__ WidomShared sh;
  __shared__ double sel_w;
  __shared__ int sel_t;
  extern __shared__ char dyn[];
  const BoxConst& bc = A.bc;
  const int T = bc.ntypes;
  // force-field rows in LDS
  float4* ffl32 = reinterpret_cast<float4*>(dyn);
  double* ffl64 = reinterpret_cast<double*>(dyn + (size_t)T * T * sizeof(float4));
  for (int i = threadIdx.x; i < T * T; i += blockDim.x) {
    ffl32[i] = make_float4(A.ff32[4 * i], A.ff32[4 * i + 1], A.ff32[4 * i + 2], 0.f);
    for (int k = 0; k < 4; ++k) ffl64[4 * i + k] = A.ff64[4 * i + k];
  }
  // this lane's framework stripe, resident in registers for the whole launch
  int n_mine = 0;
  double fx[APT], fy[APT], fz[APT], fq[APT];
  uint32_t ux[APT], uy[APT], uz[APT];
  float uq[APT];
  int ft[APT];
#pragma unroll
  for (int j = 0; j < APT; ++j) {
    const int a = j * kBlock + threadIdx.x;
    fx[j] = fy[j] = fz[j] = fq[j] = 0.0;
    ux[j] = uy[j] = uz[j] = 0u;
    uq[j] = 0.f;
    ft[j] = 0;
    if (a < A.n_fw) {
      ++n_mine;
      ft[j] = A.fw_type[a];
      if (MIXED) {
        const uint4 v = A.fw_fix[a];
        ux[j] = v.x; uy[j] = v.y; uz[j] = v.z; uq[j] = __uint_as_float(v.w);
        fq[j] = A.fw[3 * A.n_fw + a];
      } else {
        fx[j] = A.fw[a]; fy[j] = A.fw[A.n_fw + a]; fz[j] = A.fw[2 * A.n_fw + a]; fq[j] = A.fw[3 * A.n_fw + a];
      }
    }
  }
  __syncthreads();

  double blk_w = 0.0, blk_w2 = 0.0, blk_we = 0.0, blk_n = 0.0, blk_fb = 0.0, blk_ok = 0.0;
  const int K = A.K, O = A.O, chain = A.chain;
  for (int ins = blockIdx.x; ins < A.n_insert; ins += gridDim.x) {
    const uint64_t idx = (uint64_t)(A.first_index + ins);
    // 1. first-bead trial positions: diag(cell) * U^3 (get_random_trial_position, mc_widom.h:122-213)
    if (threadIdx.x < K) {
      const int k = threadIdx.x;
      double p[3];
      p[0] = bc.cell[0] * philox_uniform(A.seed, A.stream, idx, 3 * k);
      p[1] = bc.cell[4] * philox_uniform(A.seed, A.stream, idx, 3 * k + 1);
      p[2] = bc.cell[8] * philox_uniform(A.seed, A.stream, idx, 3 * k + 2);
      sh.pos[k][0] = p[0]; sh.pos[k][1] = p[1]; sh.pos[k][2] = p[2];
      if (MIXED) to_fixed(bc, p, sh.fix[k]);
    }
    __syncthreads();
    // 2. pair pass over the K trials, 3. first-bead selection
    pair_pass<MIXED, APT>(A, sh, K, 1, fx, fy, fz, fq, ux, uy, uz, uq, ft, n_mine, ffl32, ffl64);
    select_trial(A, sh, K, philox_uniform(A.seed, A.stream, idx, 3 * K), &sel_w, &sel_t);
    if (threadIdx.x == 0) {
      const int sel = sel_t;
      sh.W = sel_w;
      sh.ok = (sel >= 0 && sel_w > 1e-150) ? 1 : 0;
      sh.fb = sel;
      sh.e_fb[0] = sel >= 0 ? sh.tot[sel][0] : 0.0;
      sh.e_fb[1] = sel >= 0 ? sh.tot[sel][1] : 0.0;
      if (sel >= 0) {
        sh.mol[0][0] = sh.pos[sel][0]; sh.mol[0][1] = sh.pos[sel][1]; sh.mol[0][2] = sh.pos[sel][2];
      }
      sh.e_ch[0] = sh.e_ch[1] = 0.0;
      sh.orient = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0 && sh.ok) blk_fb += 1.0;
    if (sh.ok && chain > 0) {
      // 4. orientations of the template about the selected first bead
      //    (get_random_trial_orientation, mc_widom.h:215-303)
      if (threadIdx.x < O) {
        const int o = threadIdx.x;
        const double u = philox_uniform(A.seed, A.stream, idx, 3 * K + 1 + 3 * o);
        const double v = philox_uniform(A.seed, A.stream, idx, 3 * K + 2 + 3 * o);
        const double w = philox_uniform(A.seed, A.stream, idx, 3 * K + 3 + 3 * o);
        for (int a = 0; a < chain; ++a) {
          double vec[3] = {A.tvec[a][0], A.tvec[a][1], A.tvec[a][2]};
          quat_rotate(vec, u, v, w);
          const int s = o * chain + a;
          double p[3] = {sh.mol[0][0] + vec[0], sh.mol[0][1] + vec[1], sh.mol[0][2] + vec[2]};
          sh.pos[s][0] = p[0]; sh.pos[s][1] = p[1]; sh.pos[s][2] = p[2];
          if (MIXED) to_fixed(bc, p, sh.fix[s]);
        }
      }
      __syncthreads();
      pair_pass<MIXED, APT>(A, sh, O, chain, fx, fy, fz, fq, ux, uy, uz, uq, ft, n_mine, ffl32, ffl64);
      select_trial(A, sh, O, philox_uniform(A.seed, A.stream, idx, 3 * K + 3 * O + 1), &sel_w, &sel_t);
      if (threadIdx.x == 0) {
        const int sel = sel_t;
        if (sel < 0 || !(sh.W * sel_w > 1e-150)) {
          sh.ok = 0;
        } else {
          sh.W *= sel_w;
          sh.orient = sel;
          sh.e_ch[0] = sh.tot[sel][0]; sh.e_ch[1] = sh.tot[sel][1];
          for (int a = 0; a < chain; ++a) {
            const int s = sel * chain + a;
            sh.mol[1 + a][0] = sh.pos[s][0]; sh.mol[1 + a][1] = sh.pos[s][1]; sh.mol[1 + a][2] = sh.pos[s][2];
          }
        }
      }
      __syncthreads();
    }
    // 5. Fourier-space energy of the selected molecule, FP64 (Fourier_Ewald_Diff,
    //    Ewald_Energy_Functions.h:280-397, with Oldsize = 0)
    double ew = 0.0;
    if (sh.ok && A.charged) {
      for (int t = threadIdx.x; t < A.M * 3; t += blockDim.x) {
        const int site = t / 3, axis = t % 3;
        double s = bc.inv[axis] * sh.mol[site][0] + bc.inv[3 + axis] * sh.mol[site][1] + bc.inv[6 + axis] * sh.mol[site][2];
        s *= 2 * 3.14159265358979323846;
        double sn, cs;
        sincos(s, &sn, &cs);
        const double2 e1 = make_double2(cs, sn);
        double2* e = sh.eik[site][axis];
        e[0] = make_double2(1.0, 0.0);
        if (bc.kmax[axis] >= 1) e[1] = e1;
        for (int k = 2; k <= bc.kmax[axis]; ++k) e[k] = cmul2(e[k - 1], e1);
      }
      __syncthreads();
      double es = 0.0, ec = 0.0;
      const int kzm = bc.kmax[2], kym = bc.kmax[1];
      for (int k = threadIdx.x; k < bc.nvec; k += blockDim.x) {
        const double t = A.ktemp[k];
        if (t < 0.0) continue;
        const int kz = k % (2 * kzm + 1) - kzm;
        const int kxy = k / (2 * kzm + 1);
        const int kx = kxy / (2 * kym + 1);
        const int ky = kxy % (2 * kym + 1) - kym;
        double2 cn = make_double2(0.0, 0.0);
        for (int i = 0; i < A.M; ++i) {
          double2 ey = sh.eik[i][1][abs(ky)];
          if (ky < 0) ey.y = -ey.y;
          const double2 exy = cmul2(sh.eik[i][0][kx], ey);
          double2 ez = sh.eik[i][2][abs(kz)];
          if (kz < 0) ez.y = -ez.y;
          const double2 p = cmul2(exy, ez);
          cn.x += A.site_q[i] * p.x;
          cn.y += A.site_q[i] * p.y;
        }
        const double2 ov = A.sk_ads[k];
        const double2 nv = make_double2(ov.x + cn.x, ov.y + cn.y);
        es += t * (nv.x * nv.x + nv.y * nv.y) - t * (ov.x * ov.x + ov.y * ov.y);
        const double2 cv = A.sk_fw[k];
        ec += t * (cv.x * cn.x + cv.y * cn.y);
      }
      es = wave_sum(es);
      ec = wave_sum(ec);
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      if (lane == 0) { sh.ered[wave][0] = es; sh.ered[wave][1] = ec; }
      __syncthreads();
      const double same = sh.ered[0][0] + sh.ered[1][0] + sh.ered[2][0] + sh.ered[3][0];
      const double cross = sh.ered[0][1] + sh.ered[1][1] + sh.ered[2][1] + sh.ered[3][1];
      ew = (same - A.excl) + 2.0 * cross;
    }
    // 6. record
    if (threadIdx.x == 0) {
      double W = 0.0, vdw = 0.0, real = 0.0, etot = 0.0;
      const int ok = sh.ok;
      if (ok) {
        W = sh.W * exp(-A.beta * ew) * exp(-A.beta * A.tail_de);
        vdw = sh.e_fb[0] + sh.e_ch[0];
        real = sh.e_fb[1] + sh.e_ch[1];
        etot = vdw + real + ew + A.tail_de;
      }
      blk_w += W; blk_w2 += W * W; blk_we += W * etot; blk_n += 1.0; blk_ok += ok ? 1.0 : 0.0;
      if (A.res) {
        graspa_widom_result r;
        r.rosen = W; r.energy = etot; r.vdw = vdw; r.real = real; r.ewald = ok ? ew : 0.0;
        r.first_bead = ok ? sh.fb : -1; r.orientation = ok ? sh.orient : -1; r.ok = ok; r._pad = 0;
        A.res[ins] = r;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* o = A.wsum + kSums * blockIdx.x;
    o[0] = blk_w; o[1] = blk_w2; o[2] = blk_we; o[3] = blk_n; o[4] = blk_fb; o[5] = blk_ok;
  }
}

// Sums the per-workgroup partials in workgroup order (deterministic).
__global__ void widom_sum_kernel(double* wsum, int nblk)
{
  const int i = threadIdx.x;
  if (i < kSums) {
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += wsum[kSums * b + i];
    wsum[kSums * nblk + i] = s;
  }
}

template <bool MIXED, int N>
static int launch_one(graspa_box* b, const WidomArgs& A, int n_insert, size_t lds)
{
  int occ = 1;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, widom_kernel<MIXED, N>, kBlock, lds) != hipSuccess || occ < 1)
    occ = 1;
  const int grid = std::max(1, std::min(n_insert, occ * b->n_cu));
  hipLaunchKernelGGL((widom_kernel<MIXED, N>), dim3(grid), dim3(kBlock), lds, b->stream, A);
  if (hip_check(hipGetLastError(), "widom_kernel launch")) return -1;
  return grid;
}

template <bool MIXED>
static int launch_apt(graspa_box* b, const WidomArgs& A, int apt, int n_insert, size_t lds)
{
  if (apt <= 1) return launch_one<MIXED, 1>(b, A, n_insert, lds);
  if (apt <= 2) return launch_one<MIXED, 2>(b, A, n_insert, lds);
  if (apt <= 4) return launch_one<MIXED, 4>(b, A, n_insert, lds);
  if (apt <= 6) return launch_one<MIXED, 6>(b, A, n_insert, lds);
  if (apt <= 8) return launch_one<MIXED, 8>(b, A, n_insert, lds);
  if (apt <= 12) return launch_one<MIXED, 12>(b, A, n_insert, lds);
  if (apt <= 16) return launch_one<MIXED, 16>(b, A, n_insert, lds);
  if (apt <= 24) return launch_one<MIXED, 24>(b, A, n_insert, lds);
  fail(GRASPA_ERR_UNSUPPORTED, "framework too large for the register-resident Widom kernel (> 6144 atoms)");
  return -1;
}

int launch_widom(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                 uint64_t seed, uint32_t stream, int64_t first_index, int n_insert, graspa_widom_result* d_res,
                 graspa_widom_summary* summary)
{
  if (comp < 1 || comp >= (int)b->comps.size()) return fail(GRASPA_ERR_ARG, "widom: bad component");
  const Component& c = b->comps[comp];
  if (K < 1 || K > kMaxTrials || O < 1 || O > kMaxTrials) return fail(GRASPA_ERR_ARG, "widom: 1 <= K,O <= 32");
  if (c.molsize < 1 || c.molsize > kMaxMol) return fail(GRASPA_ERR_UNSUPPORTED, "widom: molecule larger than 8 sites");
  if (std::max(b->bc.kmax[0], std::max(b->bc.kmax[1], b->bc.kmax[2])) >= kMaxKStride)
    return fail(GRASPA_ERR_UNSUPPORTED, "widom: kmax above 15");
  if (b->bc.ntypes > kMaxTypes) return fail(GRASPA_ERR_UNSUPPORTED, "widom: too many pseudo-atom types");
  if (n_insert <= 0) {
    if (summary) *summary = graspa_widom_summary{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    return GRASPA_OK;
  }
  WidomArgs A{};
  A.bc = b->bc;
  A.fw = b->d_fw; A.fw_type = b->d_fw_type; A.fw_fix = b->d_fw_fix; A.n_fw = b->n_fw;
  A.g = b->d_g; A.g_type = b->d_g_type; A.g_molid = b->d_g_molid; A.g_fix = b->d_g_fix; A.guest_cap = b->guest_cap;
  guest_view(b, &A.gv);
  A.ff64 = b->d_ff64; A.ff32 = b->d_ff32;
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
  A.ktemp = b->d_ktemp; A.sk_ads = b->d_sk_ads; A.sk_fw = b->d_sk_fw;
  A.res = d_res;
  A.wsum = b->d_wsum;
  const int apt = (b->n_fw + kBlock - 1) / kBlock;
  const int Tp = A.bc.ntypes;
  const size_t lds = (size_t)Tp * Tp * (sizeof(float4) + 4 * sizeof(double));
  timing_begin(b, "widom");
  const int grid = (b->precision == GRASPA_MIXED) ? launch_apt<true>(b, A, apt, n_insert, lds)
                                                  : launch_apt<false>(b, A, apt, n_insert, lds);
  timing_end(b, "widom");
  if (grid < 0) return GRASPA_ERR_HIP;
  hipLaunchKernelGGL(widom_sum_kernel, dim3(1), dim3(64), 0, b->stream, b->d_wsum, grid);
  if (hip_check(hipGetLastError(), "widom_sum launch")) return GRASPA_ERR_HIP;
  if (summary) {
    double s4[kSums];
    if (hip_check(hipMemcpyAsync(s4, b->d_wsum + kSums * grid, kSums * sizeof(double), hipMemcpyDeviceToHost, b->stream),
                  "widom summary copy"))
      return GRASPA_ERR_HIP;
    if (hip_check(hipStreamSynchronize(b->stream), "widom sync")) return GRASPA_ERR_HIP;
    summary->sum_w = s4[0];
    summary->sum_w2 = s4[1];
    summary->sum_we = s4[2];
    summary->n = s4[3];
    summary->n_first_bead = s4[4];
    summary->n_ok = s4[5];
  }
  return GRASPA_OK;
}

}  // namespace graspa
