// engine_kernels.hip -- per-move kernels of the GCMC energy path on gfx950:
// probe-site pair energies (CBMC trials, moved molecules), the Fourier-space
// energy change with its structure-factor update, and the total-energy /
// structure-factor build.
//
// Per-move work is small (tens of sites against ~10^3 atoms, ~10^3 k-vectors),
// so these kernels are latency-shaped: one workgroup per probe site, sites and
// force-field rows read through the scalar path, block partials returned to the
// host in one copy.  The throughput path is the fused Widom kernel.
#include "device_math.hpp"
#include "engine.hpp"
#include "move_math.hpp"

namespace graspa {

__device__ inline void block_sum4(double v[4], double* lds /* [4*4] */)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) lds[wave * 4 + i] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < 4; ++i) {
      double s = 0.0;
      for (int w = 0; w < nw; ++w) s += lds[w * 4 + i];
      v[i] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Pair energies of one probe site (workgroup) against framework + guest atoms.
// Semantics: Calculate_Multiple_Trial_Energy_VDWReal (VDW_Coulomb.cu:1183-1352).
// out[5*s + {0..3}] = HG vdw, HG real, GG vdw, GG real; out[5*s+4] = overlap.
// ---------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(kBlock) void site_energy_kernel(
    const BoxConst bc, const SiteBatch sb, const double* __restrict__ fw, const int* __restrict__ fw_type,
    const uint4* __restrict__ fw_fix, const int n_fw, const double* __restrict__ g, const int* __restrict__ g_type,
    const int* __restrict__ g_molid, const uint4* __restrict__ g_fix, const int guest_cap, const GuestView gv,
    const double* __restrict__ ff64, const float* __restrict__ ff32, const int self_comp, const int self_molid,
    const int excl_comp, const int excl_molid, double* __restrict__ out)
{
  __shared__ double red[16];
  __shared__ int ovs;
  const int s = blockIdx.x;
  const int T = bc.ntypes;
  const double sx = sb.x[s], sy = sb.y[s], sz = sb.z[s], sq = sb.q[s];
  const int st = sb.type[s];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  int ov = 0;
  if (threadIdx.x == 0) ovs = 0;

  // framework (host-guest) then each guest component (guest-guest)
  for (int part = 0; part < gv.n_comp; ++part) {
    const int n = (part == 0) ? n_fw : gv.seg[part].n_atoms;
    const int base = (part == 0) ? 0 : gv.seg[part].base;
    for (int jj = threadIdx.x; jj < n; jj += blockDim.x) {
      const int j = base + jj;
      int ta;
      double dx, dy, dz, qa;
      if (part > 0) {
        const int mid = g_molid[j];
        if (part == excl_comp && mid == excl_molid) continue;
        if (part == self_comp && mid == self_molid) continue;
        ta = g_type[j];
        dx = g[j] - sx; dy = g[guest_cap + j] - sy; dz = g[2 * guest_cap + j] - sz; qa = g[3 * guest_cap + j];
      } else {
        ta = fw_type[j];
        dx = fw[j] - sx; dy = fw[n_fw + j] - sy; dz = fw[2 * n_fw + j] - sz; qa = fw[3 * n_fw + j];
      }
      if (!MIXED)
        pbc_fp64(dx, dy, dz, bc.cell, bc.inv, bc.cubic);
      else if (bc.cubic)
        pbc_rint<true>(dx, dy, dz, bc.cell, bc.inv);
      else
        pbc_rint<false>(dx, dy, dz, bc.cell, bc.inv);
      const double rr = dx * dx + dy * dy + dz * dz;
      double ev = 0.0, er = 0.0;
      if (rr < bc.cut_vdw2) {
        const double* f = ff64 + 4 * (ta * T + st);
        ev = MIXED ? lj_mixed(rr, f[0], f[1], f[2]) : lj_fp64(rr, f[0], f[1], f[2]);
        if (ev > bc.overlap) ov = 1;
        if (rr < 0.01) ov = 1;
      }
      if (bc.use_charges && rr < bc.cut_coul2) {
        if (MIXED) {
          er = coulomb_mixed(rr, bc.prefactor * qa * sq, bc.alpha);
        } else {
          const double r = sqrt(rr);
          const double term = qa * sq * erfc(bc.alpha * r);
          er = bc.prefactor * 1.0 * term / r;
        }
      }
      if (part == 0) { acc[0] += ev; acc[1] += er; } else { acc[2] += ev; acc[3] += er; }
    }
  }
  if (ov) atomicOr(&ovs, 1);
  block_sum4(acc, red);
  if (threadIdx.x == 0) {
    out[5 * s + 0] = acc[0];
    out[5 * s + 1] = acc[1];
    out[5 * s + 2] = acc[2];
    out[5 * s + 3] = acc[3];
    out[5 * s + 4] = (double)ovs;
  }
}

int launch_site_energies(graspa_box* b, const SiteBatch& sb, int self_comp, int self_molid, int excl_comp,
                         int excl_molid, double* d_out)
{
  if (sb.n <= 0) return GRASPA_OK;
  GuestView gv;
  guest_view(b, &gv);
  timing_begin(b, "site");
  if (b->precision == GRASPA_MIXED)
    hipLaunchKernelGGL(site_energy_kernel<true>, dim3(sb.n), dim3(kBlock), 0, b->stream, b->bc, sb, b->d_fw,
                       b->d_fw_type, b->d_fw_fix, b->n_fw, b->d_g, b->d_g_type, b->d_g_molid, b->d_g_fix,
                       b->guest_cap, gv, b->d_ff64, b->d_ff32, self_comp, self_molid, excl_comp, excl_molid, d_out);
  else
    hipLaunchKernelGGL(site_energy_kernel<false>, dim3(sb.n), dim3(kBlock), 0, b->stream, b->bc, sb, b->d_fw,
                       b->d_fw_type, b->d_fw_fix, b->n_fw, b->d_g, b->d_g_type, b->d_g_molid, b->d_g_fix,
                       b->guest_cap, gv, b->d_ff64, b->d_ff32, self_comp, self_molid, excl_comp, excl_molid, d_out);
  timing_end(b, "site");
  return hip_check(hipGetLastError(), "site_energy_kernel launch");
}

// Fourier energy change of one move (Fourier_Ewald_Diff, Ewald_Energy_Functions.h:280-397):
// one thread per k-vector; per-site axis tables in LDS.
__global__ __launch_bounds__(kBlock) void ewald_delta_kernel(const BoxConst bc, const SiteBatch sb, const int n_old,
                                                             const double* __restrict__ ktemp,
                                                             const double2* __restrict__ sk_same,
                                                             const double2* __restrict__ sk_cross,
                                                             double2* __restrict__ sk_stage, double* __restrict__ out)
{
  extern __shared__ double2 eik[];   // [n_sites][3][kstride]
  __shared__ double red[16];
  const int kstride = 1 + max(bc.kmax[0], max(bc.kmax[1], bc.kmax[2]));
  const int ns = sb.n;
  for (int t = threadIdx.x; t < ns * 3; t += blockDim.x) {
    const int site = t / 3, axis = t % 3;
    axis_table(bc, sb.x[site], sb.y[site], sb.z[site], axis, eik + (site * 3 + axis) * kstride);
  }
  __syncthreads();
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (k < bc.nvec) {
    const double t = ktemp[k];
    const double2 ov = sk_same[k];
    if (t >= 0.0) {
      const int kzm = bc.kmax[2], kym = bc.kmax[1];
      const int kz = k % (2 * kzm + 1) - kzm;
      const int kxy = k / (2 * kzm + 1);
      const int kx = kxy / (2 * kym + 1);
      const int ky = kxy % (2 * kym + 1) - kym;
      double2 co = make_double2(0.0, 0.0), cn = make_double2(0.0, 0.0);
      for (int i = 0; i < ns; ++i) {
        const double2* e = eik + i * 3 * kstride;
        double2 ey = e[kstride + abs(ky)];
        if (ky < 0) ey.y = -ey.y;
        const double2 exy = cmul2(e[kx], ey);
        double2 ez = e[2 * kstride + abs(kz)];
        if (kz < 0) ez.y = -ez.y;
        const double2 p = cmul2(exy, ez);
        const double q = sb.q[i];
        if (i < n_old) { co.x += q * p.x; co.y += q * p.y; }
        else { cn.x += q * p.x; cn.y += q * p.y; }
      }
      const double2 nv = make_double2(ov.x + cn.x - co.x, ov.y + cn.y - co.y);
      acc[0] = t * (nv.x * nv.x + nv.y * nv.y) - t * (ov.x * ov.x + ov.y * ov.y);
      const double2 cv = sk_cross[k];
      acc[1] = t * (cv.x * (cn.x - co.x) + cv.y * (cn.y - co.y));
      sk_stage[k] = nv;
    } else {
      sk_stage[k] = ov;
    }
  }
  block_sum4(acc, red);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = acc[0];
    out[2 * blockIdx.x + 1] = acc[1];
  }
}

int launch_ewald_delta(graspa_box* b, const SiteBatch& sb, int n_old, double* d_out, int* n_blocks)
{
  const int kstride = 1 + std::max(b->bc.kmax[0], std::max(b->bc.kmax[1], b->bc.kmax[2]));
  const size_t lds = (size_t)sb.n * 3 * kstride * sizeof(double2);
  if (lds > 48 * 1024) return fail(GRASPA_ERR_UNSUPPORTED, "ewald_delta: too many sites for the k range");
  const int nb = (b->nvec + kBlock - 1) / kBlock;
  *n_blocks = nb;
  timing_begin(b, "ewald");
  hipLaunchKernelGGL(ewald_delta_kernel, dim3(nb), dim3(kBlock), lds, b->stream, b->bc, sb, n_old, b->d_ktemp,
                     b->d_sk_ads, b->d_sk_fw, b->d_sk_stage, d_out);
  timing_end(b, "ewald");
  return hip_check(hipGetLastError(), "ewald_delta_kernel launch");
}

// ---------------------------------------------------------------------------
// Structure factors of the whole box (Ewald_Total, ewald_preparation.h:5-259):
// one thread per k-vector over every atom; e^{ik.r} from the same per-axis
// recurrence, built per atom in a first pass.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void atom_axis_tables(const BoxConst bc, const double* __restrict__ xyzq,
                                                           int stride, int n, int kstride, double2* __restrict__ tab)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 loc[32];
  for (int axis = 0; axis < 3; ++axis) {
    axis_table(bc, xyzq[i], xyzq[stride + i], xyzq[2 * stride + i], axis, loc);
    for (int k = 0; k <= bc.kmax[axis]; ++k) tab[((size_t)axis * kstride + k) * n + i] = loc[k];
  }
}

// Structure factors of the framework and of the adsorbates and the Fourier
// energies (Ewald_Total, ewald_preparation.h:5-259): a workgroup per
// kSkK consecutive k-vectors, its threads over the atoms (coalesced axis-table
// reads), wave sums then a fixed-order sum over the waves.
constexpr int kSkK = 4;

__global__ __launch_bounds__(kBlock) void sk_build_kernel(const BoxConst bc, const double2* __restrict__ tab_fw,
                                                          const double* __restrict__ fw, int n_fw,
                                                          const double2* __restrict__ tab_g,
                                                          const double* __restrict__ g, int guest_cap, const GuestView gv,
                                                          const double* __restrict__ ktemp, double2* __restrict__ sk_fw,
                                                          double2* __restrict__ sk_ads, double* __restrict__ out)
{
  __shared__ double red[kBlock / 64][4 * kSkK];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int kstride = 1 + max(bc.kmax[0], max(bc.kmax[1], bc.kmax[2]));
  const int kzm = bc.kmax[2], kym = bc.kmax[1];
  const int k0 = blockIdx.x * kSkK;
  int kx[kSkK], ky[kSkK], kz[kSkK];
  bool on[kSkK];
#pragma unroll
  for (int u = 0; u < kSkK; ++u) {
    const int k = k0 + u;
    on[u] = k < bc.nvec && ktemp[k] >= 0.0;
    kz[u] = k % (2 * kzm + 1) - kzm;
    const int kxy = k / (2 * kzm + 1);
    kx[u] = kxy / (2 * kym + 1);
    ky[u] = kxy % (2 * kym + 1) - kym;
  }
  // v[4u + 0..1]: framework S(k_u), v[4u + 2..3]: adsorbate S(k_u)
  double v[4 * kSkK];
#pragma unroll
  for (int i = 0; i < 4 * kSkK; ++i) v[i] = 0.0;
  for (int part = 0; part < gv.n_comp; ++part) {
    const int n = (part == 0) ? n_fw : gv.seg[part].n_atoms;
    const int base = (part == 0) ? 0 : gv.seg[part].base;
    const double2* tab = (part == 0) ? tab_fw : tab_g;
    const int tn = (part == 0) ? n_fw : guest_cap;
    const double* q = (part == 0) ? fw + 3 * n_fw : g + 3 * guest_cap;
    const int off = (part == 0) ? 0 : 2;
    for (int jj = threadIdx.x; jj < n; jj += blockDim.x) {
      const int j = base + jj;
      const double qj = q[j];
#pragma unroll
      for (int u = 0; u < kSkK; ++u) {
        if (!on[u]) continue;
        double2 ey = tab[((size_t)1 * kstride + abs(ky[u])) * tn + j];
        if (ky[u] < 0) ey.y = -ey.y;
        const double2 exy = cmul2(tab[((size_t)0 * kstride + kx[u]) * tn + j], ey);
        double2 ez = tab[((size_t)2 * kstride + abs(kz[u])) * tn + j];
        if (kz[u] < 0) ez.y = -ez.y;
        const double2 p = cmul2(exy, ez);
        v[4 * u + off] += qj * p.x;
        v[4 * u + off + 1] += qj * p.y;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 4 * kSkK; ++i) v[i] = wave_sum(v[i]);
  if (lane == 0)
    for (int i = 0; i < 4 * kSkK; ++i) red[wave][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int u = 0; u < kSkK; ++u) {
      const int k = k0 + u;
      if (k >= bc.nvec) break;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
        for (int i = 0; i < 4; ++i) s[i] += red[w][4 * u + i];
      const double2 sf = make_double2(s[0], s[1]), sa = make_double2(s[2], s[3]);
      sk_fw[k] = sf;
      sk_ads[k] = sa;
      if (on[u]) {
        const double t = ktemp[k];
        acc[0] += t * (sa.x * sa.x + sa.y * sa.y);           // guest-guest
        acc[1] += t * (sf.x * sf.x + sf.y * sf.y);           // host-host
        acc[2] += t * (sf.x * sa.x + sf.y * sa.y) * 2.0;     // host-guest
      }
    }
    out[3 * blockIdx.x] = acc[0];
    out[3 * blockIdx.x + 1] = acc[1];
    out[3 * blockIdx.x + 2] = acc[2];
  }
}

int launch_ewald_build(graspa_box* b, double* d_out, int* n_blocks)
{
  const int kstride = 1 + std::max(b->bc.kmax[0], std::max(b->bc.kmax[1], b->bc.kmax[2]));
  if (kstride > 32) return fail(GRASPA_ERR_UNSUPPORTED, "kmax above 31");
  GuestView gv;
  guest_view(b, &gv);
  double2 *tab_fw = nullptr, *tab_g = nullptr;
  const size_t nf = (size_t)std::max(b->n_fw, 1), ng = (size_t)std::max(b->guest_cap, 1);
  if (hip_check(hipMalloc(&tab_fw, 3 * kstride * nf * sizeof(double2)), "hipMalloc eik framework")) return GRASPA_ERR_HIP;
  if (hip_check(hipMalloc(&tab_g, 3 * kstride * ng * sizeof(double2)), "hipMalloc eik guests")) {
    hipFree(tab_fw);
    return GRASPA_ERR_HIP;
  }
  if (b->n_fw > 0)
    hipLaunchKernelGGL(atom_axis_tables, dim3((b->n_fw + kBlock - 1) / kBlock), dim3(kBlock), 0, b->stream, b->bc,
                       b->d_fw, b->n_fw, b->n_fw, kstride, tab_fw);
  if (b->guest_cap > 0)
    hipLaunchKernelGGL(atom_axis_tables, dim3((b->guest_cap + kBlock - 1) / kBlock), dim3(kBlock), 0, b->stream,
                       b->bc, b->d_g, b->guest_cap, b->guest_cap, kstride, tab_g);
  const int nb = (b->nvec + kSkK - 1) / kSkK;
  *n_blocks = nb;
  hipLaunchKernelGGL(sk_build_kernel, dim3(nb), dim3(kBlock), 0, b->stream, b->bc, tab_fw, b->d_fw, b->n_fw, tab_g,
                     b->d_g, b->guest_cap, gv, b->d_ktemp, b->d_sk_fw, b->d_sk_ads, d_out);
  int rc = hip_check(hipGetLastError(), "sk_build launch");
  hipStreamSynchronize(b->stream);
  hipFree(tab_fw);
  hipFree(tab_g);
  return rc;
}

// ---------------------------------------------------------------------------
// Total VDW + real-space energy (VDWReal_Total_CPU, VDW_Coulomb.cu:34-226):
// one thread per guest atom against the framework (weight 1: the reference
// counts each host-guest pair from both sides at 1/2) and the other guest
// molecules (weight 1/2).  Host-host pairs of the single rigid framework
// molecule are all excluded by the reference's same-molecule rule.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void total_pairs_kernel(const BoxConst bc, const double* __restrict__ fw,
                                                             const int* __restrict__ fw_type, const int n_fw,
                                                             const double* __restrict__ g,
                                                             const int* __restrict__ g_type,
                                                             const int* __restrict__ g_molid,
                                                             const int* __restrict__ g_comp, const int guest_cap,
                                                             const GuestView gv, const double* __restrict__ ff64,
                                                             double* __restrict__ out)
{
  __shared__ double red[16];
  const int T = bc.ntypes;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  // map i to a live guest atom
  int ci = -1, ia = -1;
  {
    int off = i;
    for (int c = 1; c < gv.n_comp; ++c) {
      if (off < gv.seg[c].n_atoms) { ci = c; ia = gv.seg[c].base + off; break; }
      off -= gv.seg[c].n_atoms;
    }
  }
  if (ci > 0) {
    const double xi = g[ia], yi = g[guest_cap + ia], zi = g[2 * guest_cap + ia], qi = g[3 * guest_cap + ia];
    const int ti = g_type[ia], mi = g_molid[ia];
    for (int part = 0; part < gv.n_comp; ++part) {
      const int n = (part == 0) ? n_fw : gv.seg[part].n_atoms;
      const int base = (part == 0) ? 0 : gv.seg[part].base;
      for (int jj = 0; jj < n; ++jj) {
        const int j = base + jj;
        double dx, dy, dz, qj;
        int tj;
        if (part == 0) {
          dx = xi - fw[j]; dy = yi - fw[n_fw + j]; dz = zi - fw[2 * n_fw + j]; qj = fw[3 * n_fw + j]; tj = fw_type[j];
        } else {
          if (part == ci && g_molid[j] == mi) continue;
          dx = xi - g[j]; dy = yi - g[guest_cap + j]; dz = zi - g[2 * guest_cap + j]; qj = g[3 * guest_cap + j];
          tj = g_type[j];
        }
        pbc_fp64(dx, dy, dz, bc.cell, bc.inv, bc.cubic);
        const double rr = dx * dx + dy * dy + dz * dz;
        const double w = (part == 0) ? 1.0 : 0.5;
        if (rr < bc.cut_vdw2) {
          const double* f = ff64 + 4 * (ti * T + tj);
          const double e = lj_fp64(rr, f[0], f[1], f[2]);
          acc[part == 0 ? 0 : 2] += w * e;
        }
        if (bc.use_charges && rr < bc.cut_coul2) {
          const double r = sqrt(rr);
          acc[part == 0 ? 1 : 3] += w * (bc.prefactor * (qi * qj * erfc(bc.alpha * r)) / r);
        }
      }
    }
  }
  block_sum4(acc, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) out[4 * blockIdx.x + k] = acc[k];
}

int launch_total_pairs(graspa_box* b, double* d_out, int* n_blocks)
{
  GuestView gv;
  guest_view(b, &gv);
  const int n = std::max(gv.total, 1);
  const int nb = (n + kBlock - 1) / kBlock;
  *n_blocks = nb;
  hipLaunchKernelGGL(total_pairs_kernel, dim3(nb), dim3(kBlock), 0, b->stream, b->bc, b->d_fw, b->d_fw_type, b->n_fw,
                     b->d_g, b->d_g_type, b->d_g_molid, b->d_g_comp, b->guest_cap, gv, b->d_ff64, d_out);
  return hip_check(hipGetLastError(), "total_pairs launch");
}

}  // namespace graspa
