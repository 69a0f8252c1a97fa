// batch_mc.hip -- device-resident GCMC of many independent boxes in one
// persistent launch on gfx950: one workgroup (8 waves) owns one box and runs
// its Monte Carlo steps with no host round trip.
//
// One step, as RunMoves (axpy.cu:102-298) does it:
//   selection of component, molecule and move (Select_Box_Component_Molecule,
//   axpy.cu:74-101);
//   translation / rotation (SingleBodyMove, mc_single_particle.h:10-314;
//   get_new_position, mc_utilities.h:485-606): pair energy change of the
//   moved molecule, overlap on the new position, Ewald change, Metropolis;
//   insertion / deletion / Widom (Insertion_Body / Deletion_Body,
//   mc_swap_utilities.h:3-225, with the CBMC trial functions of mc_widom.h:14-614
//   and GetPrefactor, mc_utilities.h:315-351): K first-bead trials, selection,
//   O orientations of the template about the selected bead, selection, Ewald
//   and tail terms, Rosenbluth weight, acceptance;
//   commit (mc_utilities.h:163-292, 357-417): positions, the last molecule
//   into a deleted molecule's hole, and the structure factors by flipping the
//   box's double buffer (Update_Vector_Ewald, Ewald_Energy_Functions.h:423-433);
// cycles of max(20, N_molecules + 1) steps with the max-step update every 500
// cycles (axpy.cu:378-400, 580-591).
//
// Work inside a step: the pair energies of up to 64 probe sites (trial sites,
// or the old and new sites of a moved molecule) against every framework and
// guest atom run over all 512 lanes (lanes over atoms, up to 8 sites per pass
// held in registers, FP64 sums reduced by a transposing butterfly across the
// wave, then in wave order across the workgroup, so every sum is
// deterministic); the Ewald change runs one lane per active k-vector with
// e^{ik.r} axis tables of the moved sites in LDS; selections, the Metropolis
// test and bookkeeping run on wave 0.  The force-field rows and, in the fast
// form, the real-space Coulomb table are staged in LDS once per launch.
//
// Random numbers: move t of a box uses philox_uniform(seed, stream, t, d) with
// the draw layout documented in graspa_engine.h, so a box's trajectory does
// not depend on the other boxes of the batch or on the launch boundaries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_math.hpp"
#include "engine.hpp"
#include "move_math.hpp"

namespace graspa {

constexpr int kBT = 512;               // threads of a box workgroup
constexpr int kBW = kBT / 64;          // waves
constexpr int kBMaxTrial = 32;         // K and O
constexpr int kBMaxMol = 8;            // atoms per molecule
constexpr int kBMaxSites = 64;         // probe sites of one pair pass
constexpr int kBGroup = 8;             // sites per pass over the atoms
// neighbour list of a pair pass (cubic cells): per wave a segment of the
// framework atoms and one of the guest atoms within the pass's reach
constexpr int kLocFw = 256;            // framework entries per wave
constexpr int kLocG = 128;             // guest entries per wave
constexpr int kBDrawTrial = 16;        // first trial draw
constexpr int kBMaxDraws = kBDrawTrial + 9 * kBMaxTrial;   // K trials, O orientations, O retrace orientations

enum { kMvTrans = GRASPA_MOVE_TRANSLATION, kMvRot = GRASPA_MOVE_ROTATION, kMvIns = GRASPA_MOVE_INSERTION,
       kMvDel = GRASPA_MOVE_DELETION, kMvWidom = GRASPA_MOVE_WIDOM, kMvNone = GRASPA_MOVE_NONE,
       kMvReins = GRASPA_MOVE_REINSERTION, kMvIdent = GRASPA_MOVE_IDENTITY_SWAP };
// draws of an identity swap: old species, new species, old molecule
enum { kBDrawIdOld = 11, kBDrawIdNew = 12, kBDrawIdMol = 13 };

// Per-species set-up of one box (the Components fields RunMoves reads).
struct BSpecies {
  double p_trans, p_rot, p_widom, p_swap;     // cumulative
  double p_reins;                             // cumulative, between p_widom and p_swap
  double p_ident;                             // cumulative, between p_reins and p_swap
  double mol_fraction, fugacity, ideal_rosen, excl;   // excl = ExclusionIntra + ExclusionAtom
  int base, molsize, capacity, charged;
  int pseudo[GRASPA_MAX_PSEUDO];
  double q[kBMaxMol];               // charges and types of the species' atoms (the same in every
  int type[kBMaxMol];               // molecule)
  const double* chain;              // flexible chain model (kChainDoubles), null for a rigid molecule
};

// Read-only description of one box.
struct BBox {
  BoxConst bc;
  double volume;
  const double* fw;                 // SoA x, y, z, q with stride n_fw
  const int* fw_type;
  int n_fw;
  int guest_cap;
  double* g;                        // SoA x, y, z, q with stride guest_cap
  int* g_type;
  int* g_molid;
  int* g_comp;
  const double* ff64;               // {4 eps, sigma^2, shift, eps} per type pair
  const double* coul;
  int coul_n;
  int kstride;
  const int4* kact;                 // active k-vectors {kx, ky, kz, index}
  const double* ktemp_act;
  int n_active;
  int has_tail;
  const double2* skfw;              // framework S(k) of the active k-vectors
  double2* skc[2];                  // adsorbate S(k) of the active k-vectors, double buffer
  const uint4* ctab;                // overlap-screen cells: cap slots of {fx, fy, fz, type}
  int cell_n[3];
  int cell_cap;                     // 0: no cell grid
  const double* tail;               // [T*T]
  const int* use_tail;
  int n_comp, n_valid;
  int valid[kMaxComps];             // species with SwapProbability > 1e-10
  BSpecies sp[kMaxComps];
  int K, O, bias, use_max_step;
  int max_step, trace_cap;
  int cyc_cap, n_block;             // n_block: blocks of a production run
  double beta, pressure;
  uint64_t seed;
  uint32_t stream, _pad2;
  graspa_trace* trace;
  float* cyc;                       // per-cycle ticks (ring of cyc_cap)
  graspa_block_stats* blk;          // block statistics of the production run
};

// Mutable state of one box; held in LDS during a launch.
struct BState {
  int64_t move, cycle, target;
  int step, steps;
  int sk_cur, error, production, _pad;
  int n_mol[kMaxComps];
  int64_t total_mol;
  int64_t n_pseudo[GRASPA_MAX_PSEUDO];
  double max_trans[kMaxComps][3], max_rot[kMaxComps][3];
  int64_t cnt[kMaxComps][GRASPA_N_COUNTERS];
  int64_t upd_tot[kMaxComps][2], upd_acc[kMaxComps][2];
  double widom_w[kMaxComps], widom_n[kMaxComps];
  double running[10];
  double load_sum[kMaxComps], load_sum2[kMaxComps];
  int64_t load_n;
  uint64_t cycle_t0;
  int64_t trace_n;
  int64_t cyc_first;                // first cycle of the current run
  int64_t prod_first, blk_size;     // first cycle and block length of the production run
  uint64_t busy;                    // device-clock ticks spent in launches
};
static_assert(sizeof(BState) % 8 == 0, "BState is copied as 64-bit words");
static_assert(kMaxComps == GRASPA_MAX_COMPS, "block statistics are sized for kMaxComps components");
static_assert(kBT >= 10 + 3 * (kMaxComps - 1) + (kMaxComps - 1) * (kMaxComps - 1) + 1,
              "one thread per block sum of a cycle");
static_assert(sizeof(BBox) % 8 == 0, "BBox is copied as 64-bit words");

struct BSite {
  double x, y, z, q;
  int type, _p;
};

// LDS workspace of one step.
struct BMove {
  BSite site[kBMaxSites];
  uint4 pfix[kBMaxSites];           // sites as fixed-point cell fractions (screen)
  unsigned char idx[kBMaxSites];    // sites of the current pair pass
  double e[kBMaxSites][4];          // hg vdw, hg real, gg vdw, gg real per site
  double part[2][kBW][2 * kBGroup];
  double ud[kBMaxDraws];
  double lw[kBMaxTrial], shf[kBMaxTrial], ew[kBMaxTrial], t4[kBMaxTrial][4];
  double tmpl[kMaxComps][kBMaxMol][3];   // slot 0 of each species, read at the start of a move
  double molp[kBMaxMol][3], molq[kBMaxMol];
  int molt[kBMaxMol];
  double ewp[kBW][2];
  double en[10];                    // graspa_energy of the move
  double W, pacc, same, cross;
  double Wo, stored, eo[4];          // reinsertion: retrace weight, other first-bead weights, old energies
  unsigned long long ovl, smask;
  int type, tmove, comp, mol, M, start, charged, n_eval;
  int fail, accept, overlap;
  int nc, qo, qn, _pad;             // identity swap: new species, old / new species charged
  // guest atoms mirrored in LDS for the pair passes: gl_n atoms in SoA
  // (x, y, z, q, then types); component c's atoms [base, base + gl_cap[c])
  // at gl_off[c] (gl_cap 0: not mirrored)
  int gl_n, gl_off[kMaxComps], gl_cap[kMaxComps];
  // neighbour list of the current pair pass: centre and squared reach,
  // per-wave entry counts (framework, guests) and their prefix sums
  double lc[3], lr2;
  int loc_on, loc_over;
  int lcnt[2][kBW], lpre[2][kBW + 1];
};

// Keeps the LDS mirror of guest atom j of species c equal to the store.
__device__ inline void mirror_put(const BBox& B, const BMove& L, double* gml, const int c, const int j, const double x,
                                  const double y, const double z, const double q, const int t)
{
  const int k = j - B.sp[c].base;
  if (gml && k < L.gl_cap[c]) {
    const int n = L.gl_n, m = L.gl_off[c] + k;
    gml[m] = x; gml[n + m] = y; gml[2 * n + m] = z; gml[3 * n + m] = q;
    reinterpret_cast<int*>(gml + 4 * n)[m] = t;
  }
}

__device__ inline void mirror_pos(const BBox& B, const BMove& L, double* gml, const int c, const int j, const double x,
                                  const double y, const double z)
{
  const int k = j - B.sp[c].base;
  if (gml && k < L.gl_cap[c]) {
    const int n = L.gl_n, m = L.gl_off[c] + k;
    gml[m] = x; gml[n + m] = y; gml[2 * n + m] = z;
  }
}

// ---------------------------------------------------------------------------
// pair energies of the probe sites L.site[0 .. n_sites)
// ---------------------------------------------------------------------------

// Wave sums of the 2 * kBGroup values of every lane (LJ and real-space sums
// of the pass's sites): halving exchanges, each lane keeping the half its lane
// bit selects, until each group of 64 / NV lanes holds one value, then plain
// exchanges inside the group.  The first lane of group g writes value g.
__device__ inline void wave_sum_group(const double av[kBGroup], const double ar[kBGroup], double* out)
{
  constexpr int NV = 2 * kBGroup;
  static_assert(NV >= 2 && NV <= 64 && (NV & (NV - 1)) == 0, "power-of-two value count");
  const int lane = threadIdx.x & 63;
  double v[NV];
#pragma unroll
  for (int j = 0; j < kBGroup; ++j) { v[2 * j] = av[j]; v[2 * j + 1] = ar[j]; }
  int idx = 0;
#pragma unroll
  for (int n = NV, m = 32; n > 1; n >>= 1, m >>= 1) {
    const bool hi = (lane & m) != 0;
#pragma unroll
    for (int i = 0; i < n / 2; ++i) {
      const double keep = hi ? v[n / 2 + i] : v[i];
      const double send = hi ? v[i] : v[n / 2 + i];
      v[i] = keep + __shfl_xor(send, m, 64);
    }
    if (hi) idx += n / 2;
  }
#pragma unroll
  for (int m = 64 / NV / 2; m >= 1; m >>= 1) v[0] += __shfl_xor(v[0], m, 64);
  if ((lane & (64 / NV - 1)) == 0) out[idx] = v[0];
}

// OR over the 64 lanes (every lane active), uniform result: the steps of
// wave_sum on both 32-bit halves (data-parallel moves within the SIMD, no LDS
// round trips as the lane shuffles take), lane 63 read back.
__device__ inline uint32_t wave_or32(uint32_t v)
{
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xf, 0xf, false);    // quad_perm [1,0,3,2]
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xf, 0xf, false);    // quad_perm [2,3,0,1]
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x124, 0xf, 0xf, false);   // row_ror:4
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x128, 0xf, 0xf, false);   // row_ror:8
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);   // row_bcast:15
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);   // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ inline unsigned long long wave_or64(unsigned long long v)
{
  const uint32_t lo = wave_or32((uint32_t)v), hi = wave_or32((uint32_t)(v >> 32));
  return ((unsigned long long)hi << 32) | lo;
}

// Semantics of one pair: Calculate_Multiple_Trial_Energy_VDWReal
// (VDW_Coulomb.cu:1183-1352) and the single-body kernel (VDW_Coulomb.cu:626-841):
// LJ inside the VDW cutoff with the overlap flag (E > OverlapCriteria or
// r^2 < 0.01), erfc real space inside the Coulomb cutoff.
template <bool MIXED, bool CUBIC>
__device__ inline void pair_site(const BoxConst& bc, const double* __restrict__ ffl, const double2* __restrict__ ct,
                                 const int T, const BSite& s, const double ax, const double ay,
                                 const double az, const double aq, const int at, double& ev, double& er, uint32_t& ov,
                                 const uint32_t bit)
{
  double dx = ax - s.x, dy = ay - s.y, dz = az - s.z;
  if (MIXED) pbc_rint<CUBIC>(dx, dy, dz, bc.cell, bc.inv);
  else pbc_fp64(dx, dy, dz, bc.cell, bc.inv, CUBIC ? 1 : 0);
  const double rr = dx * dx + dy * dy + dz * dz;
  if (rr < bc.cut_vdw2) {
    const double* f = ffl + 4 * (at * T + s.type);
    const double e = MIXED ? lj_mixed(rr, f[0], f[1], f[2]) : lj_fp64(rr, f[0], f[1], f[2]);
    ev += e;
    if (e > bc.overlap || rr < 0.01) ov |= bit;
  }
  if (bc.use_charges && rr < bc.cut_coul2) {
    if (MIXED) {
      er += (bc.prefactor * aq * s.q) * coul_table(ct, rr);
    } else {
      const double r = sqrt(rr);
      const double term = aq * s.q * erfc(bc.alpha * r);
      er += bc.prefactor * 1.0 * term / r;
    }
  }
}

// Energies of the sites L.idx[0 .. n_sites) into L.e[site] and their overlap
// flags into L.ovl (bit site), against the framework (host-guest) and every
// guest atom except the index range [ex_lo, ex_hi) (the molecule itself).
// The sites form G groups of at most kBGroup; the 8 waves are dealt out to
// the groups (kBW / G waves each), so all groups run in one pass over the
// atoms with every lane busy, and one reduction finishes the pass.
// Block of the current cycle in the production run (Gather_Averages_Types,
// print_statistics.cuh:646-655: cycle / block size; the remainder goes to the
// last block).
__device__ inline int block_of(const BState& S, const int n_block)
{
  const int64_t b = (S.cycle - S.prod_first) / S.blk_size;
  return (int)(b < n_block - 1 ? b : n_block - 1);
}

// Block sums of one production cycle (axpy.cu:420-453): the running energy
// change per term, and per adsorbate N, N^2, N times the running change
// without its host-host terms, N_c N_d, and the cycle count; one thread per
// sum, each sum a single read-modify-write.
__device__ void block_sums(const BBox& B, const BState& S, const int bk, const int tid)
{
  graspa_block_stats& st = *B.blk;
  const int na = B.n_comp - 1;
  int t = tid;
  if (t < 10) { st.e[bk][t] += S.running[t]; return; }
  t -= 10;
  if (t < 3 * na) {
    const int c = 1 + t / 3, k = t % 3;
    const double nm = (double)S.n_mol[c];
    if (k == 0) st.n[bk][c] += nm;
    else if (k == 1) st.n2[bk][c] += nm * nm;
    else {
      const double ads = S.running[1] + S.running[2] + S.running[4] + S.running[5] + S.running[7] +
                         S.running[8] + S.running[9];
      st.en[bk][c] += ads * nm;
    }
    return;
  }
  t -= 3 * na;
  if (t < na * na) {
    const int c = 1 + t / na, d = 1 + t % na;
    st.nn[bk][c][d] += (double)S.n_mol[c] * (double)S.n_mol[d];
    return;
  }
  t -= na * na;
  if (t == 0) st.cycles[bk] += 1;
}

template <bool MIXED, bool CUBIC>
__device__ void pair_pass(const BBox& B, const BState& S, BMove& L, const double* __restrict__ ffl,
                          const double2* __restrict__ ct, const double* __restrict__ fwl,
                          const double* __restrict__ gml, int* __restrict__ loc, const int n_sites, const int ex_lo,
                          const int ex_hi)
{
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const BoxConst& bc = B.bc;
  const int T = bc.ntypes;
  // Neighbour list (cubic cells): when the sites lie within R of their
  // centre c and the sphere of radius r_cut + R about c is well inside the
  // cell, only atoms with |a - c| < r_cut + R (minimum image) can be within
  // the cutoff of a site (triangle inequality), so the pass runs over those.
  if (CUBIC && loc) {
    if (wave == 0) {
      // centre and radius of the sites (minimum image about site 0), one
      // lane per site
      const BSite& s0 = L.site[L.idx[0]];
      double dx = 0.0, dy = 0.0, dz = 0.0;
      if (lane < n_sites) {
        const BSite& sk = L.site[L.idx[lane]];
        dx = sk.x - s0.x; dy = sk.y - s0.y; dz = sk.z - s0.z;
        pbc_rint<true>(dx, dy, dz, bc.cell, bc.inv);
      }
      const double inv_n = 1.0 / (double)n_sites;
      const double cx = wave_sum(dx) * inv_n, cy = wave_sum(dy) * inv_n, cz = wave_sum(dz) * inv_n;
      const double ex = dx - cx, ey = dy - cy, ez = dz - cz;
      const double r2 = wave_max(lane < n_sites ? ex * ex + ey * ey + ez * ez : 0.0);
      if (lane == 0) {
        const double rcut = sqrt(bc.use_charges ? fmax(bc.cut_vdw2, bc.cut_coul2) : bc.cut_vdw2);
        const double reach = rcut + sqrt(r2) + 1e-6;
        const double lmin = fmin(bc.cell[0], fmin(bc.cell[4], bc.cell[8]));
        L.lc[0] = s0.x + cx; L.lc[1] = s0.y + cy; L.lc[2] = s0.z + cz;
        L.lr2 = reach * reach;
        // worth it when the sphere holds under half the cell's volume
        L.loc_on = (4.18879020478639 * reach * reach * reach < 0.5 * B.volume && reach < 0.5 * lmin) ? 1 : 0;
        L.loc_over = 0;
      }
    }
    __syncthreads();
    if (L.loc_on) {
      // wave w filters a contiguous share of the atoms in index order into
      // its own segments (deterministic order)
      const unsigned long long lt = (1ull << lane) - 1ull;
      const double cx = L.lc[0], cy = L.lc[1], cz = L.lc[2], lr2 = L.lr2;
      const int nf = B.n_fw, cap = B.guest_cap;
      const int share = (nf + kBW - 1) / kBW;
      int cnt = 0;
      for (int i0 = wave * share; i0 < min(nf, (wave + 1) * share); i0 += 64) {
        const int i = i0 + lane;
        bool in = false;
        if (i < min(nf, (wave + 1) * share)) {
          double dx = (fwl ? fwl[i] : B.fw[i]) - cx, dy = (fwl ? fwl[nf + i] : B.fw[nf + i]) - cy;
          double dz = (fwl ? fwl[2 * nf + i] : B.fw[2 * nf + i]) - cz;
          pbc_rint<true>(dx, dy, dz, bc.cell, bc.inv);
          in = dx * dx + dy * dy + dz * dz < lr2;
        }
        const unsigned long long m = __ballot(in);
        const int k = cnt + __popcll(m & lt);
        if (in && k < kLocFw) loc[wave * kLocFw + k] = i;
        cnt += __popcll(m);
      }
      int gcnt = 0;
      for (int c = 1; c < B.n_comp; ++c) {
        const int lo = B.sp[c].base, hi = lo + S.n_mol[c] * B.sp[c].molsize;
        const int gs = (hi - lo + kBW - 1) / kBW;
        const int a0 = lo + wave * gs, a1 = min(hi, lo + (wave + 1) * gs);
        for (int i0 = a0; i0 < a1; i0 += 64) {
          const int i = i0 + lane;
          bool in = false;
          if (i < a1 && !(i >= ex_lo && i < ex_hi)) {
            double dx = B.g[i] - cx, dy = B.g[cap + i] - cy, dz = B.g[2 * cap + i] - cz;
            pbc_rint<true>(dx, dy, dz, bc.cell, bc.inv);
            in = dx * dx + dy * dy + dz * dz < lr2;
          }
          const unsigned long long m = __ballot(in);
          const int k = gcnt + __popcll(m & lt);
          if (in && k < kLocG) loc[kBW * kLocFw + wave * kLocG + k] = i;
          gcnt += __popcll(m);
        }
      }
      if (lane == 0) {
        L.lcnt[0][wave] = cnt;
        L.lcnt[1][wave] = gcnt;
        if (cnt > kLocFw || gcnt > kLocG) L.loc_over = 1;
      }
      __syncthreads();
      if (tid < 2) {
        int a = 0;
        L.lpre[tid][0] = 0;
        for (int w = 0; w < kBW; ++w) L.lpre[tid][w + 1] = (a += L.lcnt[tid][w]);
      }
      __syncthreads();
    }
  }
  const bool use_loc = CUBIC && loc && L.loc_on && !L.loc_over;
  // G groups of at most kBGroup sites, the sites and the waves split evenly
  // between them: group g holds sites [g n / G, (g+1) n / G) and waves
  // [g kBW / G, (g+1) kBW / G), so every wave is busy
  const int ngrp = (n_sites + kBGroup - 1) / kBGroup;    // <= kBW
  int grp = 0;
  while (grp + 1 < ngrp && wave >= (grp + 1) * kBW / ngrp) ++grp;
  const int w0 = grp * kBW / ngrp, w1 = (grp + 1) * kBW / ngrp;
  const int g0 = grp * n_sites / ngrp;
  const int cnt = (grp + 1) * n_sites / ngrp - g0;
  const int r = wave - w0;
  const int first = r * 64 + lane, stride = (w1 - w0) * 64;
  const int nf = B.n_fw, cap = B.guest_cap;
  uint32_t ovb = 0;
#pragma unroll 1
  for (int part = 0; part < 2; ++part) {
    double av[kBGroup], ar[kBGroup];
#pragma unroll
    for (int j = 0; j < kBGroup; ++j) { av[j] = 0.0; ar[j] = 0.0; }
    if (use_loc) {
      // the neighbour list: entry k of the part lies in the segment of the
      // last wave whose prefix is <= k
      const int* pre = L.lpre[part];
      const int n_in = pre[kBW];
      const int* seg = loc + (part == 0 ? 0 : kBW * kLocFw);
      const int segcap = part == 0 ? kLocFw : kLocG;
      const int* __restrict__ ftl = fwl ? reinterpret_cast<const int*>(fwl + 4 * nf) : nullptr;
      const int gn = L.gl_n;
      const int* __restrict__ gtl = gml ? reinterpret_cast<const int*>(gml + 4 * gn) : nullptr;
#pragma unroll 1
      for (int k = first; k < n_in; k += stride) {
        int w = 0;
        while (w + 1 < kBW && pre[w + 1] <= k) ++w;
        const int i = seg[w * segcap + (k - pre[w])];
        double ax, ay, az, aq;
        int at;
        if (part == 0) {
          if (fwl) { ax = fwl[i]; ay = fwl[nf + i]; az = fwl[2 * nf + i]; aq = fwl[3 * nf + i]; at = ftl[i]; }
          else { ax = B.fw[i]; ay = B.fw[nf + i]; az = B.fw[2 * nf + i]; aq = B.fw[3 * nf + i]; at = B.fw_type[i]; }
        } else {
          // guest i of species c: from the mirror when the species fits there
          int c = 1;
          while (c + 1 < B.n_comp && i >= B.sp[c + 1].base) ++c;
          const int lo = B.sp[c].base;
          if (gml && S.n_mol[c] * B.sp[c].molsize <= L.gl_cap[c]) {
            const int m = i - lo + L.gl_off[c];
            ax = gml[m]; ay = gml[gn + m]; az = gml[2 * gn + m]; aq = gml[3 * gn + m]; at = gtl[m];
          } else {
            ax = B.g[i]; ay = B.g[cap + i]; az = B.g[2 * cap + i]; aq = B.g[3 * cap + i]; at = B.g_type[i];
          }
        }
#pragma unroll
        for (int j = 0; j < kBGroup; ++j)
          if (j < cnt)
            pair_site<MIXED, CUBIC>(bc, ffl, ct, T, L.site[L.idx[g0 + j]], ax, ay, az, aq, at, av[j], ar[j],
                                    ovb, 1u << j);
      }
    } else if (part == 0 && fwl) {
      // framework staged in LDS for the launch: SoA x, y, z, q, then types
      const int* __restrict__ ftl = reinterpret_cast<const int*>(fwl + 4 * nf);
#pragma unroll 1
      for (int i = first; i < nf; i += stride) {
        const double ax = fwl[i], ay = fwl[nf + i], az = fwl[2 * nf + i], aq = fwl[3 * nf + i];
        const int at = ftl[i];
#pragma unroll
        for (int j = 0; j < kBGroup; ++j)
          if (j < cnt) pair_site<MIXED, CUBIC>(bc, ffl, ct, T, L.site[L.idx[g0 + j]], ax, ay, az, aq, at, av[j], ar[j],
                                          ovb, 1u << j);
      }
    } else if (part == 0) {
#pragma unroll 1
      for (int i = first; i < nf; i += stride) {
        const double ax = B.fw[i], ay = B.fw[nf + i], az = B.fw[2 * nf + i], aq = B.fw[3 * nf + i];
        const int at = B.fw_type[i];
#pragma unroll
        for (int j = 0; j < kBGroup; ++j)
          if (j < cnt) pair_site<MIXED, CUBIC>(bc, ffl, ct, T, L.site[L.idx[g0 + j]], ax, ay, az, aq, at, av[j], ar[j],
                                          ovb, 1u << j);
      }
    } else {
      for (int c = 1; c < B.n_comp; ++c) {
        const int lo = B.sp[c].base, n_at = S.n_mol[c] * B.sp[c].molsize, hi = lo + n_at;
        // from the LDS mirror when the species' atoms all fit in it
        const bool lds = gml && n_at <= L.gl_cap[c];
        const int gn = L.gl_n, go = L.gl_off[c] - lo;
        const int* __restrict__ gtl = reinterpret_cast<const int*>(gml + 4 * gn);
#pragma unroll 1
        for (int i = lo + first; i < hi; i += stride) {
          if (i >= ex_lo && i < ex_hi) continue;
          double ax, ay, az, aq;
          int at;
          if (lds) {
            const int m = i + go;
            ax = gml[m]; ay = gml[gn + m]; az = gml[2 * gn + m]; aq = gml[3 * gn + m];
            at = gtl[m];
          } else {
            ax = B.g[i]; ay = B.g[cap + i]; az = B.g[2 * cap + i]; aq = B.g[3 * cap + i];
            at = B.g_type[i];
          }
#pragma unroll
          for (int j = 0; j < kBGroup; ++j)
            if (j < cnt)
              pair_site<MIXED, CUBIC>(bc, ffl, ct, T, L.site[L.idx[g0 + j]], ax, ay, az, aq, at, av[j], ar[j],
                                          ovb, 1u << j);
        }
      }
    }
    wave_sum_group(av, ar, L.part[part][wave]);
  }
  unsigned long long ovl = 0ull;
#pragma unroll
  for (int j = 0; j < kBGroup; ++j)
    if (j < cnt && ((ovb >> j) & 1u)) ovl |= 1ull << L.idx[g0 + j];
  ovl = wave_or64(ovl);
  if (lane == 0 && ovl) atomicOr(&L.ovl, ovl);
  __syncthreads();
  if (tid < 4 * n_sites) {
    // site slot -> its group (the group boundaries as above), summed over the
    // group's waves in wave order
    const int slot = tid >> 2, t = tid & 3, part = t >> 1, k = t & 1;
    int g = 0;
    while (g + 1 < ngrp && slot >= (g + 1) * n_sites / ngrp) ++g;
    const int j = slot - g * n_sites / ngrp;
    double s = 0.0;
    for (int w = g * kBW / ngrp; w < (g + 1) * kBW / ngrp; ++w) s += L.part[part][w][2 * j + k];
    L.e[L.idx[slot]][t] = s;
  }
  __syncthreads();
}

// Overlap screen of sites [first, first + n): FP32 squared distances from
// fixed-point cell fractions to the framework atoms of the 27 grid cells
// around each site (the cells are at least as wide as the largest overlap
// distance, so they hold every framework atom that can overlap the site)
// against thresholds r_ov^2 lowered by 1e-5 relative: a hit is a pair the
// exact evaluation flags as an overlap in either precision form.  Hits go to
// L.smask (bit site).  Needs a cell grid (B.cell_cap > 0).
template <bool CUBIC>
__device__ void screen_pass(const BBox& B, BMove& L, const float* __restrict__ thr, const int first, const int n)
{
  const int tid = threadIdx.x, lane = tid & 63;
  const int T = B.bc.ntypes, C = B.cell_cap;
  const int n0 = B.cell_n[0], n1 = B.cell_n[1], n2 = B.cell_n[2];
  if (tid < n) {
    const BSite& st = L.site[first + tid];
    const double p[3] = {st.x, st.y, st.z};
    L.pfix[first + tid] = to_fixed4(B.bc, p);
  }
  if (tid == 0) L.smask = 0ull;
  __syncthreads();
  const int per = 27 * C, total = n * per;
  unsigned long long m = 0ull;
  for (int t = tid; t < total; t += kBT) {
    const int k = t / per, r = t - k * per, nb = r / C, slot = r - nb * C;
    const int site = first + k;
    const uint4 P = L.pfix[site];
    int x = (int)(((uint64_t)P.x * (uint64_t)n0) >> 32) + nb % 3 - 1;
    int y = (int)(((uint64_t)P.y * (uint64_t)n1) >> 32) + (nb / 3) % 3 - 1;
    int z = (int)(((uint64_t)P.z * (uint64_t)n2) >> 32) + nb / 9 - 1;
    x = x < 0 ? x + n0 : (x >= n0 ? x - n0 : x);
    y = y < 0 ? y + n1 : (y >= n1 ? y - n1 : y);
    z = z < 0 ? z + n2 : (z >= n2 ? z - n2 : z);
    const uint4 at = B.ctab[(size_t)(x + n0 * (y + n1 * z)) * C + slot];
    if (at.w != ~0u) {
      const float th = thr[L.site[site].type * T + (int)at.w];
      if (r2_fixed32<CUBIC>(at.x, at.y, at.z, P.x, P.y, P.z, B.bc.h32) < th) m |= 1ull << site;
    }
  }
  m = wave_or64(m);
  if (lane == 0 && m) atomicOr(&L.smask, m);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Ewald: Fourier energy change of removing sites [0, n_old) and adding sites
// [n_old, n_old + n_new) of L.site (GPU_EwaldDifference_General,
// Ewald_Energy_Functions.h:438-580 with Fourier_Ewald_Diff :280-397): raw
// same-type and cross sums into L.same / L.cross; the updated adsorbate S(k)
// goes to the box's other buffer.
// ---------------------------------------------------------------------------
__device__ void ewald_pass(const BBox& B, const BState& S, BMove& L, double2* __restrict__ eik, const int n_old,
                           const int n_new)
{
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ns = n_old + n_new, ks = B.kstride;
  for (int t = tid; t < ns * 3; t += kBT) {
    const int site = t / 3, axis = t - 3 * site;
    axis_table(B.bc, L.site[site].x, L.site[site].y, L.site[site].z, axis, eik + (site * 3 + axis) * ks);
  }
  __syncthreads();
  const double2* __restrict__ cur = B.skc[S.sk_cur];
  double2* __restrict__ stg = B.skc[S.sk_cur ^ 1];
  double a0 = 0.0, a1 = 0.0;
  // kEwU k-vectors per thread per step, their loads issued together (one
  // workgroup per box leaves few waves to hide memory latency); a thread
  // still visits its k-vectors in ascending order, so the sums are the same
  constexpr int kEwU = 8;
  const int na = B.n_active;
  for (int j0 = tid; j0 < na; j0 += kBT * kEwU) {
    int4 kv[kEwU];
    double tk[kEwU];
    double2 ov[kEwU], cv[kEwU];
#pragma unroll
    for (int u = 0; u < kEwU; ++u) {
      const int j = j0 + u * kBT;
      if (j < na) {
        kv[u] = B.kact[j];
        tk[u] = B.ktemp_act[j];
        ov[u] = cur[j];
        cv[u] = B.skfw[j];
      }
    }
#pragma unroll
    for (int u = 0; u < kEwU; ++u) {
      const int j = j0 + u * kBT;
      if (j >= na) break;
      double2 co = make_double2(0.0, 0.0), cn = make_double2(0.0, 0.0);
      for (int i = 0; i < ns; ++i) {
        const double2* e = eik + i * 3 * ks;
        double2 ey = e[ks + abs(kv[u].y)];
        if (kv[u].y < 0) ey.y = -ey.y;
        double2 ez = e[2 * ks + abs(kv[u].z)];
        if (kv[u].z < 0) ez.y = -ez.y;
        const double2 p = cmul2(cmul2(e[kv[u].x], ey), ez);
        const double q = L.site[i].q;
        if (i < n_old) { co.x += q * p.x; co.y += q * p.y; }
        else { cn.x += q * p.x; cn.y += q * p.y; }
      }
      const double2 nv = make_double2(ov[u].x + cn.x - co.x, ov[u].y + cn.y - co.y);
      a0 += tk[u] * (nv.x * nv.x + nv.y * nv.y) - tk[u] * (ov[u].x * ov[u].x + ov[u].y * ov[u].y);
      a1 += tk[u] * (cv[u].x * (cn.x - co.x) + cv[u].y * (cn.y - co.y));
      stg[j] = nv;
    }
  }
  a0 = wave_sum(a0);
  a1 = wave_sum(a1);
  if (lane == 0) { L.ewp[wave][0] = a0; L.ewp[wave][1] = a1; }
  __syncthreads();
  if (tid == 0) {
    double s0 = 0.0, s1 = 0.0;
    for (int w = 0; w < kBW; ++w) { s0 += L.ewp[w][0]; s1 += L.ewp[w][1]; }
    L.same = s0;
    L.cross = s1;
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// CBMC selection on wave 0 (SelectTrialPosition, mc_widom.h:14-39;
// CBMC_FirstBead_Finish :305-383; Widom_Move_Chain_PARTIAL :574-614): trial
// sums L.t4[0..n) with overlap bits ovm.  Survivors in trial order, log-sum-exp
// shift, cumulative pick with u (insertion) or the first survivor
// (deletion), Rosenbluth sum over n, the real-space correction when the bias
// is LJ only.  Returns the selected trial (-1 on failure) and multiplies *W.
// ---------------------------------------------------------------------------
__device__ int select_wave0(const BBox& B, BMove& L, const int n, const uint32_t ovm, const double u, const bool del,
                            const bool chain, double* W, double* others = nullptr)
{
  const int lane = threadIdx.x & 63;
  const bool alive = lane < n && !((ovm >> lane) & 1u);
  double lw = -INFINITY, er = 0.0;
  if (lane < n) er = L.t4[lane][1] + L.t4[lane][3];
  if (alive) {
    double tot = L.t4[lane][0] + L.t4[lane][2];
    if (B.bias) tot += er;
    lw = -B.beta * tot;
  }
  const double mx = wave_max(lw);
  const double shf = alive ? exp(lw - mx) : -1.0;   // shifted factors; -1 marks an overlap
  const double ew = alive ? exp(lw) : 0.0;
  // ordered sums over the survivors (uniform values through lane reads)
  int nk = 0, first = -1;
  double ssum = 0.0, rb = 0.0;
  for (int t = 0; t < n; ++t) {
    const double sv = readlane_d(shf, t);
    if (sv >= 0.0) {
      if (first < 0) first = t;
      ssum += sv;
      rb += readlane_d(ew, t);
      ++nk;
    }
  }
  int sel = -1;
  bool fail = false;
  if (!del) {
    if (nk == 0) {
      fail = true;
    } else {
      const double ws = u * ssum;
      int t = first;
      double cumw = readlane_d(shf, t);
      int left = nk - 1;
      while (cumw < ws && left > 0) {
        ++t;
        while (readlane_d(shf, t) < 0.0) ++t;
        cumw += readlane_d(shf, t);
        --left;
      }
      sel = t;
      if (rb < 1e-150) fail = true;
    }
  } else {
    sel = first;
    if (chain && nk == 0) fail = true;
  }
  if (!fail) {
    // the weights of the survivors other than the selected one (StoredR of a
    // reinsertion, CBMC_FirstBead_Finish mc_widom.h:365)
    if (others && sel >= 0) *others = rb - readlane_d(ew, sel);
    double w = rb / (double)n;
    if (sel >= 0 && !B.bias) w *= exp(-B.beta * readlane_d(er, sel));
    w *= *W;
    if (w <= 1e-150) fail = true;
    *W = w;
  }
  return fail ? -1 : sel;
}

// TailCorrectionDifference (TailCorrection_Energy_Functions.h:36-82)
__device__ double tail_delta(const BBox& B, const BState& S, const int comp, const int sign)
{
  const int T = B.bc.ntypes;
  double e = 0.0;
  for (int i = 0; i < T && i < GRASPA_MAX_PSEUDO; ++i) {
    const int64_t di = (int64_t)B.sp[comp].pseudo[i] * sign;
    for (int j = i; j < T && j < GRASPA_MAX_PSEUDO; ++j) {
      const int64_t dj = (int64_t)B.sp[comp].pseudo[j] * sign;
      const int ij = i * T + j;
      if (!B.use_tail[ij]) continue;
      const int64_t dn = S.n_pseudo[i] * dj + S.n_pseudo[j] * di + di * dj;
      double de = B.tail[ij] * (double)dn;
      if (i != j) de *= 2.0;
      e += de;
    }
  }
  return e / B.volume;
}

// TailCorrectionIdentitySwap (TailCorrection_Energy_Functions.h:85-108): the
// pseudo-atom change is one molecule of species nc in, one of oc out
__device__ double tail_identity(const BBox& B, const BState& S, const int nc, const int oc)
{
  const int T = B.bc.ntypes;
  double e = 0.0;
  for (int i = 0; i < T && i < GRASPA_MAX_PSEUDO; ++i) {
    const int64_t di = (int64_t)B.sp[nc].pseudo[i] - B.sp[oc].pseudo[i];
    for (int j = i; j < T && j < GRASPA_MAX_PSEUDO; ++j) {
      const int64_t dj = (int64_t)B.sp[nc].pseudo[j] - B.sp[oc].pseudo[j];
      const int ij = i * T + j;
      if (!B.use_tail[ij]) continue;
      const int64_t dn = S.n_pseudo[i] * dj + S.n_pseudo[j] * di + di * dj;
      double de = B.tail[ij] * (double)dn;
      if (i != j) de *= 2.0;
      e += de;
    }
  }
  return e / B.volume;
}

__device__ inline void add_energy(double* run, const double* en)
{
  for (int i = 0; i < 10; ++i) run[i] += en[i];
}
// graspa_energy order: hh_vdw, hg_vdw, gg_vdw, hh_real, hg_real, gg_real,
// hh_ewald, hg_ewald, gg_ewald, tail
enum { kHGV = 1, kGGV = 2, kHGR = 4, kGGR = 5, kHGE = 7, kGGE = 8, kTail = 9 };

// ---------------------------------------------------------------------------
// the moves
// ---------------------------------------------------------------------------
template <bool MIXED, bool CUBIC>
__device__ void single_body(const BBox& B, BState& S, BMove& L, const double* ffl, const float* thr, const double2* ct,
                            const double* fwl, double* gml, int* loc, double2* eik)
{
  const int tid = threadIdx.x;
  const int M = L.M, start = L.start, cap = B.guest_cap, comp = L.comp;
  if (tid < M) {
    const int j = start + tid;
    BSite s;
    s.x = B.g[j]; s.y = B.g[cap + j]; s.z = B.g[2 * cap + j];
    s.q = B.sp[comp].q[tid];
    s.type = B.sp[comp].type[tid]; s._p = 0;
    L.site[tid] = s;
  }
  __syncthreads();
  if (tid < M) {
    const BSite o = L.site[tid];
    BSite s = o;
    if (L.type == kMvTrans) {
      s.x = o.x + S.max_trans[comp][0] * 2.0 * (L.ud[8] - 0.5);
      s.y = o.y + S.max_trans[comp][1] * 2.0 * (L.ud[9] - 0.5);
      s.z = o.z + S.max_trans[comp][2] * 2.0 * (L.ud[10] - 0.5);
    } else {
      const BSite a0 = L.site[0];
      double p[3] = {o.x - a0.x, o.y - a0.y, o.z - a0.z};
      rotate_axis(p, S.max_rot[comp][0] * 2.0 * (L.ud[8] - 0.5), 1.0, 0.0, 0.0);
      rotate_axis(p, S.max_rot[comp][1] * 2.0 * (L.ud[9] - 0.5), 0.0, 1.0, 0.0);
      rotate_axis(p, S.max_rot[comp][2] * 2.0 * (L.ud[10] - 0.5), 0.0, 0.0, 1.0);
      s.x = p[0] + a0.x; s.y = p[1] + a0.y; s.z = p[2] + a0.z;
    }
    L.site[M + tid] = s;
  }
  __syncthreads();
  if (B.cell_cap > 0) {
    // a certain overlap of the new position rejects the move (no energies needed)
    screen_pass<CUBIC>(B, L, thr, M, M);
    if (L.smask) {
      if (tid == 0) L.overlap = 1;
      return;
    }
  }
  if (tid < 2 * M) L.idx[tid] = (unsigned char)tid;
  if (tid == 0) L.ovl = 0ull;
  __syncthreads();
  pair_pass<MIXED, CUBIC>(B, S, L, ffl, ct, fwl, gml, loc, 2 * M, start, start + M);
  if (tid == 0) {
    const unsigned long long newmask = ((1ull << M) - 1ull) << M;
    L.overlap = (L.ovl & newmask) ? 1 : 0;
    if (!L.overlap) {
      for (int a = 0; a < M; ++a) {
        L.en[kHGV] += L.e[M + a][0] - L.e[a][0];
        L.en[kHGR] += L.e[M + a][1] - L.e[a][1];
        L.en[kGGV] += L.e[M + a][2] - L.e[a][2];
        L.en[kGGR] += L.e[M + a][3] - L.e[a][3];
      }
    }
  }
  __syncthreads();
  if (L.overlap) return;
  if (L.charged) {
    ewald_pass(B, S, L, eik, M, M);
    if (tid == 0) {
      L.en[kGGE] = L.same;
      L.en[kHGE] = 2.0 * L.cross;
    }
  }
  if (tid == 0) {
    double tot = 0.0;
    for (int i = 0; i < 10; ++i) tot += L.en[i];
    L.pacc = 1.0 * exp(-B.beta * tot);
    L.accept = (L.ud[4] < L.pacc) ? 1 : 0;
    const int k = (L.type == kMvTrans) ? 0 : 1;
    if (L.accept) {
      S.cnt[comp][2 * k + 1]++;
      S.upd_acc[comp][k]++;
      add_energy(S.running, L.en);
      if (L.charged) S.sk_cur ^= 1;
    }
  }
  __syncthreads();
  if (L.accept && tid < M) {
    const int j = start + tid;
    const BSite s = L.site[M + tid];
    B.g[j] = s.x; B.g[cap + j] = s.y; B.g[2 * cap + j] = s.z;
    mirror_pos(B, L, gml, comp, j, s.x, s.y, s.z);
  }
}

// Survivors of the screen among sites [0, n) in groups of g consecutive sites
// (a trial is dead if any of its sites is screened out) become the pass's
// site list; the screened sites' overlap bits are set.  Thread 0.
__device__ inline void list_survivors(BMove& L, const int n, const int g)
{
  const unsigned long long gm = (g >= 64) ? ~0ull : ((1ull << g) - 1ull);
  int k = 0;
  for (int t0 = 0; t0 < n; t0 += g) {
    if ((L.smask >> t0) & gm) continue;
    for (int a = 0; a < g; ++a) L.idx[k++] = (unsigned char)(t0 + a);
  }
  L.n_eval = k;
  L.ovl = L.smask;
}

template <bool MIXED, bool CUBIC>
__device__ void cbmc(const BBox& B, BState& S, BMove& L, const double* ffl, const float* thr, const double2* ct,
                     const double* fwl, double* gml, int* loc, double2* eik)
{
  const int tid = threadIdx.x;
  const int M = L.M, comp = L.comp, cap = B.guest_cap;
  const int K = B.K, O = B.O, chain = M - 1;
  const bool del = (L.type == kMvDel);
  const BSpecies& sp = B.sp[comp];
  const int ex_lo = del ? L.start : 0, ex_hi = del ? L.start + M : 0;
  // first-bead trials (get_random_trial_position, mc_widom.h:122-213)
  if (tid < K) {
    BSite s;
    if (del && tid == 0) {
      s.x = B.g[L.start]; s.y = B.g[cap + L.start]; s.z = B.g[2 * cap + L.start];
    } else {
      s.x = B.bc.cell[0] * L.ud[kBDrawTrial + 3 * tid];
      s.y = B.bc.cell[4] * L.ud[kBDrawTrial + 3 * tid + 1];
      s.z = B.bc.cell[8] * L.ud[kBDrawTrial + 3 * tid + 2];
    }
    s.q = sp.q[0];
    s.type = sp.type[0];
    s._p = 0;
    L.site[tid] = s;
  }
  __syncthreads();
  if (B.cell_cap > 0) screen_pass<CUBIC>(B, L, thr, 0, K);
  else if (tid == 0) L.smask = 0ull;
  if (tid == 0) list_survivors(L, K, 1);
  __syncthreads();
  if (L.n_eval > 0) pair_pass<MIXED, CUBIC>(B, S, L, ffl, ct, fwl, gml, loc, L.n_eval, ex_lo, ex_hi);
  if (tid < K) {
#pragma unroll
    for (int t = 0; t < 4; ++t) L.t4[tid][t] = L.e[tid][t];
  }
  __syncthreads();
  if (tid < 64) {
    double W = 1.0;
    const int sel = select_wave0(B, L, K, (uint32_t)L.ovl, L.ud[5], del, false, &W);
    if (tid == 0) {
      if (sel < 0) {
        L.fail = 1;
      } else {
        L.W = W;
        L.en[kHGV] = L.t4[sel][0]; L.en[kHGR] = L.t4[sel][1];
        L.en[kGGV] = L.t4[sel][2]; L.en[kGGR] = L.t4[sel][3];
        L.molp[0][0] = L.site[sel].x; L.molp[0][1] = L.site[sel].y; L.molp[0][2] = L.site[sel].z;
        L.molq[0] = L.site[sel].q;
        L.molt[0] = L.site[sel].type;
      }
    }
  }
  __syncthreads();
  if (L.fail) return;
  if (chain > 0 && sp.chain) {
    // flexible chain, bead by bead: O trial positions of bead i from the
    // bonded densities given the beads placed (a deletion retraces its own
    // bead as trial 0), screened, evaluated and selected on the Boltzmann
    // factor of their external energy
    for (int i = 1; i < M; ++i) {
      if (tid < O) {
        BSite s;
        if (del && tid == 0) {
          const int j = L.start + i;
          s.x = B.g[j]; s.y = B.g[cap + j]; s.z = B.g[2 * cap + j];
        } else {
          double p[3];
          flex_trial(L.molp, i, sp.chain, B.beta, B.seed, B.stream, (uint64_t)S.move, tid, p);
          s.x = p[0]; s.y = p[1]; s.z = p[2];
        }
        s.q = sp.q[i];
        s.type = sp.type[i];
        s._p = 0;
        L.site[tid] = s;
      }
      __syncthreads();
      if (B.cell_cap > 0) screen_pass<CUBIC>(B, L, thr, 0, O);
      else if (tid == 0) L.smask = 0ull;
      if (tid == 0) list_survivors(L, O, 1);
      __syncthreads();
      if (L.n_eval > 0) pair_pass<MIXED, CUBIC>(B, S, L, ffl, ct, fwl, gml, loc, L.n_eval, ex_lo, ex_hi);
      if (tid < O) {
#pragma unroll
        for (int t = 0; t < 4; ++t) L.t4[tid][t] = L.e[tid][t];
      }
      __syncthreads();
      if (tid < 64) {
        double W = L.W;
        const double u = philox_uniform(B.seed, B.stream, (uint64_t)S.move, kFlexDraw - 8 + i);
        const int sel = select_wave0(B, L, O, (uint32_t)L.ovl, u, del, true, &W);
        if (tid == 0) {
          if (sel < 0) {
            L.fail = 1;
          } else {
            L.W = W;
            L.en[kHGV] += L.t4[sel][0]; L.en[kHGR] += L.t4[sel][1];
            L.en[kGGV] += L.t4[sel][2]; L.en[kGGR] += L.t4[sel][3];
            L.molp[i][0] = L.site[sel].x; L.molp[i][1] = L.site[sel].y; L.molp[i][2] = L.site[sel].z;
            L.molq[i] = L.site[sel].q;
            L.molt[i] = L.site[sel].type;
          }
        }
      }
      __syncthreads();
      if (L.fail) return;
    }
  } else if (chain > 0) {
    // orientations (get_random_trial_orientation, mc_widom.h:215-303): the
    // template vectors of the component's first slot as it is now, rotated
    for (int t = tid; t < O * chain; t += kBT) {
      const int o = t / chain, a = t - o * chain;
      BSite s;
      if (del && o == 0) {
        const int j = L.start + 1 + a;
        s.x = B.g[j]; s.y = B.g[cap + j]; s.z = B.g[2 * cap + j];
      } else {
        const double* t0 = L.tmpl[comp][0];
        const double* t1 = L.tmpl[comp][1 + a];
        double v[3] = {t1[0] - t0[0], t1[1] - t0[1], t1[2] - t0[2]};
        const int d = kBDrawTrial + 3 * K + 3 * o;
        quat_rotate(v, L.ud[d], L.ud[d + 1], L.ud[d + 2]);
        s.x = L.molp[0][0] + v[0]; s.y = L.molp[0][1] + v[1]; s.z = L.molp[0][2] + v[2];
      }
      s.q = sp.q[1 + a];
      s.type = sp.type[1 + a];
      s._p = 0;
      L.site[t] = s;
    }
    __syncthreads();
    if (B.cell_cap > 0) screen_pass<CUBIC>(B, L, thr, 0, O * chain);
    else if (tid == 0) L.smask = 0ull;
    if (tid == 0) list_survivors(L, O * chain, chain);
    __syncthreads();
    if (L.n_eval > 0) pair_pass<MIXED, CUBIC>(B, S, L, ffl, ct, fwl, gml, loc, L.n_eval, ex_lo, ex_hi);
    if (tid < O) {
      double e4[4] = {0.0, 0.0, 0.0, 0.0};
      for (int a = 0; a < chain; ++a)
        for (int t = 0; t < 4; ++t) e4[t] += L.e[tid * chain + a][t];
      for (int t = 0; t < 4; ++t) L.t4[tid][t] = e4[t];
    }
    uint32_t ovo = 0;
    const unsigned long long amask = (chain >= 64) ? ~0ull : ((1ull << chain) - 1ull);
    for (int o = 0; o < O; ++o)
      if ((L.ovl >> (o * chain)) & amask) ovo |= 1u << o;
    __syncthreads();
    if (tid < 64) {
      double W = L.W;
      const int sel = select_wave0(B, L, O, ovo, L.ud[6], del, true, &W);
      if (tid == 0) {
        if (sel < 0) {
          L.fail = 1;
        } else {
          L.W = W;
          L.en[kHGV] += L.t4[sel][0]; L.en[kHGR] += L.t4[sel][1];
          L.en[kGGV] += L.t4[sel][2]; L.en[kGGR] += L.t4[sel][3];
          for (int a = 0; a < chain; ++a) {
            const BSite& s = L.site[sel * chain + a];
            L.molp[1 + a][0] = s.x; L.molp[1 + a][1] = s.y; L.molp[1 + a][2] = s.z;
            L.molq[1 + a] = s.q;
            L.molt[1 + a] = s.type;
          }
        }
      }
    }
    __syncthreads();
    if (L.fail) return;
  }
  // Ewald of the whole molecule: the new one, or the real one for a deletion
  if (L.charged) {
    if (tid < M) {
      BSite s;
      if (del) {
        const int j = L.start + tid;
        s.x = B.g[j]; s.y = B.g[cap + j]; s.z = B.g[2 * cap + j]; s.q = sp.q[tid]; s.type = sp.type[tid];
      } else {
        s.x = L.molp[tid][0]; s.y = L.molp[tid][1]; s.z = L.molp[tid][2]; s.q = L.molq[tid]; s.type = L.molt[tid];
      }
      s._p = 0;
      L.site[tid] = s;
    }
    __syncthreads();
    ewald_pass(B, S, L, eik, del ? M : 0, del ? 0 : M);
  }
  if (tid == 0) {
    double W = L.W;
    const double N = (double)S.n_mol[comp];
    const double bz = B.beta * sp.mol_fraction * B.pressure * sp.fugacity * B.volume;
    const double pref = del ? N / bz : bz / (1.0 + N);
    if (L.charged) {
      double same = L.same;
      if (!del) {
        same -= sp.excl * 1.0;
        L.en[kGGE] = same;
        L.en[kHGE] = 2.0 * L.cross;
        W *= exp(-B.beta * (same + 2.0 * L.cross));
      } else {
        same -= sp.excl * (0.0 - 1.0);
        W /= exp(-B.beta * (same + 2.0 * L.cross));
        L.en[kGGE] = -1.0 * same;
        L.en[kHGE] = -1.0 * (2.0 * L.cross);
      }
    }
    if (B.has_tail) {
      const double te = tail_delta(B, S, comp, del ? -1 : 1);
      if (!del) { W *= exp(-B.beta * te); L.en[kTail] = te; }
      else { W /= exp(-B.beta * te); L.en[kTail] = -te; }
    }
    L.W = W;
    if (L.type == kMvWidom) {
      if (S.production) {
        S.widom_w[comp] += W;
        S.widom_n[comp] += 1.0;
        const int bk = block_of(S, B.n_block);
        B.blk->widom_w[bk][comp] += W;
        B.blk->widom_n[bk][comp] += 1.0;
      }
    } else {
      double pacc = del ? pref * sp.ideal_rosen / W : pref * W / sp.ideal_rosen;
      if (!del && W <= 1e-150) pacc = 0.0;
      L.pacc = pacc;
      int acc = (L.ud[4] < pacc) ? 1 : 0;
      if (acc && !del && S.n_mol[comp] + 1 > sp.capacity) {
        S.error = GRASPA_ERR_CAPACITY;
        acc = 0;
      }
      L.accept = acc;
      if (acc) {
        if (!del) {
          S.total_mol++;
          for (int t = 0; t < GRASPA_MAX_PSEUDO; ++t) S.n_pseudo[t] += sp.pseudo[t];
          S.cnt[comp][5]++;
        } else {
          S.total_mol--;
          for (int t = 0; t < GRASPA_MAX_PSEUDO; ++t) S.n_pseudo[t] -= sp.pseudo[t];
          S.cnt[comp][7]++;
          for (int i = 0; i < 10; ++i) L.en[i] = -L.en[i];
        }
        add_energy(S.running, L.en);
        if (L.charged) S.sk_cur ^= 1;
      }
    }
  }
  __syncthreads();
  if (L.accept) {
    const int N = S.n_mol[comp];
    if (!del) {
      if (tid < M) {
        const int j = sp.base + N * M + tid;
        B.g[j] = L.molp[tid][0]; B.g[cap + j] = L.molp[tid][1]; B.g[2 * cap + j] = L.molp[tid][2];
        B.g[3 * cap + j] = L.molq[tid];
        B.g_type[j] = L.molt[tid];
        B.g_molid[j] = N;
        B.g_comp[j] = comp;
        mirror_put(B, L, gml, comp, j, L.molp[tid][0], L.molp[tid][1], L.molp[tid][2], L.molq[tid], L.molt[tid]);
      }
    } else {
      // the last molecule moves into the hole, whose MolIDs stay
      // (Update_deletion_data_Parallel, mc_utilities.h:163-189)
      const int last = sp.base + (N - 1) * M;
      if (L.start != last && tid < M) {
        const int j = L.start + tid, k = last + tid;
        const double x = B.g[k], y = B.g[cap + k], z = B.g[2 * cap + k], q = B.g[3 * cap + k];
        const int t = B.g_type[k];
        B.g[j] = x; B.g[cap + j] = y; B.g[2 * cap + j] = z;
        B.g[3 * cap + j] = q;
        B.g_type[j] = t;
        mirror_put(B, L, gml, comp, j, x, y, z, q, t);
      }
    }
    __syncthreads();
    if (tid == 0) S.n_mol[comp] += del ? -1 : 1;
  }
}

// Pair energies of the sites L.site[0 .. n) into L.t4 per trial (chain
// sites of a trial summed) and the trials' overlap bits: the overlap screen,
// then the pair pass over the survivors; the molecule [ex_lo, ex_hi) is
// excluded.
template <bool MIXED, bool CUBIC>
__device__ uint32_t trial_energies(const BBox& B, BState& S, BMove& L, const double* ffl, const float* thr,
                                   const double2* ct, const double* fwl, double* gml, int* loc,
                                   const int n_trials, const int chain, const int ex_lo, const int ex_hi)
{
  const int tid = threadIdx.x;
  const int ns = n_trials * chain;
  if (B.cell_cap > 0) screen_pass<CUBIC>(B, L, thr, 0, ns);
  else if (tid == 0) L.smask = 0ull;
  if (tid == 0) list_survivors(L, ns, chain);
  __syncthreads();
  if (L.n_eval > 0) pair_pass<MIXED, CUBIC>(B, S, L, ffl, ct, fwl, gml, loc, L.n_eval, ex_lo, ex_hi);
  if (tid < n_trials) {
    double e4[4] = {0.0, 0.0, 0.0, 0.0};
    for (int a = 0; a < chain; ++a)
      for (int t = 0; t < 4; ++t) e4[t] += L.e[tid * chain + a][t];
    for (int t = 0; t < 4; ++t) L.t4[tid][t] = e4[t];
  }
  uint32_t ov = 0;
  const unsigned long long amask = (chain >= 64) ? ~0ull : ((1ull << chain) - 1ull);
  for (int o = 0; o < n_trials; ++o)
    if ((L.ovl >> (o * chain)) & amask) ov |= 1u << o;
  __syncthreads();
  return ov;
}

// Retrace weight of trials 0 .. n on thread 0: the ordered sum of the
// survivors' Boltzmann factors plus `others`, divided by `div`; the first
// survivor (SelectedTrial 0 of the retrace, mc_widom.h:350-356, 588-604)
// gives the energies added to L.eo and the real-space correction when the
// bias is LJ only.
__device__ double retrace_weight(const BBox& B, BMove& L, const int n, const uint32_t ov, const double others,
                                 const int div)
{
  double rb = others;
  int first = -1;
  for (int t = 0; t < n; ++t) {
    if ((ov >> t) & 1u) continue;
    double tot = L.t4[t][0] + L.t4[t][2];
    if (B.bias) tot += L.t4[t][1] + L.t4[t][3];
    rb += exp(-B.beta * tot);
    if (first < 0) first = t;
  }
  double w = rb / (double)div;
  if (first >= 0) {
    if (!B.bias) w *= exp(-B.beta * (L.t4[first][1] + L.t4[first][3]));
    for (int t = 0; t < 4; ++t) L.eo[t] += L.t4[first][t];
  }
  return w;
}

// CBMC reinsertion (ReinsertionMove::Run, move_struct.h:170-407): a new
// configuration grown as an insertion (K random first beads, O orientations
// of the template about the selected one), the old one retraced (its own
// first bead, whose weight joins the growth's other first-bead weights, and
// O orientations with trial 0 its own), energies without the molecule
// itself; W_new times the Ewald factor of moving the molecule; Pacc =
// W_new / W_old.  Draws: trial k at kBDrawTrial + 3k, orientation o at
// kBDrawTrial + 3K + 3o, retrace orientation o at kBDrawTrial + 3K + 3O + 3o.
template <bool MIXED, bool CUBIC>
__device__ void reinsertion(const BBox& B, BState& S, BMove& L, const double* ffl, const float* thr,
                            const double2* ct, const double* fwl, double* gml, int* loc, double2* eik)
{
  const int tid = threadIdx.x;
  const int M = L.M, comp = L.comp, cap = B.guest_cap;
  const int K = B.K, O = B.O, chain = M - 1;
  const BSpecies& sp = B.sp[comp];
  const int ex_lo = L.start, ex_hi = L.start + M;
  // growth: first bead
  if (tid < K) {
    BSite s;
    s.x = B.bc.cell[0] * L.ud[kBDrawTrial + 3 * tid];
    s.y = B.bc.cell[4] * L.ud[kBDrawTrial + 3 * tid + 1];
    s.z = B.bc.cell[8] * L.ud[kBDrawTrial + 3 * tid + 2];
    s.q = sp.q[0];
    s.type = sp.type[0];
    s._p = 0;
    L.site[tid] = s;
  }
  __syncthreads();
  uint32_t ov = trial_energies<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, K, 1, ex_lo, ex_hi);
  if (tid < 64) {
    double W = 1.0, others = 0.0;
    const int sel = select_wave0(B, L, K, ov, L.ud[5], false, false, &W, &others);
    if (tid == 0) {
      if (sel < 0) {
        L.fail = 1;
      } else {
        L.W = W;
        L.stored = others;
        L.en[kHGV] = L.t4[sel][0]; L.en[kHGR] = L.t4[sel][1];
        L.en[kGGV] = L.t4[sel][2]; L.en[kGGR] = L.t4[sel][3];
        L.molp[0][0] = L.site[sel].x; L.molp[0][1] = L.site[sel].y; L.molp[0][2] = L.site[sel].z;
      }
    }
  }
  __syncthreads();
  if (L.fail) return;
  // growth: orientations about the selected first bead
  if (chain > 0) {
    for (int t = tid; t < O * chain; t += kBT) {
      const int o = t / chain, a = t - o * chain;
      const double* t0 = L.tmpl[comp][0];
      const double* t1 = L.tmpl[comp][1 + a];
      double v[3] = {t1[0] - t0[0], t1[1] - t0[1], t1[2] - t0[2]};
      const int d = kBDrawTrial + 3 * K + 3 * o;
      quat_rotate(v, L.ud[d], L.ud[d + 1], L.ud[d + 2]);
      BSite s;
      s.x = L.molp[0][0] + v[0]; s.y = L.molp[0][1] + v[1]; s.z = L.molp[0][2] + v[2];
      s.q = sp.q[1 + a];
      s.type = sp.type[1 + a];
      s._p = 0;
      L.site[t] = s;
    }
    __syncthreads();
    ov = trial_energies<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, O, chain, ex_lo, ex_hi);
    if (tid < 64) {
      double W = L.W;
      const int sel = select_wave0(B, L, O, ov, L.ud[6], false, true, &W);
      if (tid == 0) {
        if (sel < 0) {
          L.fail = 1;
        } else {
          L.W = W;
          L.en[kHGV] += L.t4[sel][0]; L.en[kHGR] += L.t4[sel][1];
          L.en[kGGV] += L.t4[sel][2]; L.en[kGGR] += L.t4[sel][3];
          for (int a = 0; a < chain; ++a) {
            const BSite& s = L.site[sel * chain + a];
            L.molp[1 + a][0] = s.x; L.molp[1 + a][1] = s.y; L.molp[1 + a][2] = s.z;
          }
        }
      }
    }
    __syncthreads();
    if (L.fail) return;
  }
  // retrace: the molecule's own first bead
  if (tid == 0) {
    BSite s;
    s.x = B.g[L.start]; s.y = B.g[cap + L.start]; s.z = B.g[2 * cap + L.start];
    s.q = sp.q[0];
    s.type = sp.type[0];
    s._p = 0;
    L.site[0] = s;
    for (int t = 0; t < 4; ++t) L.eo[t] = 0.0;
  }
  __syncthreads();
  ov = trial_energies<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, 1, 1, ex_lo, ex_hi);
  if (tid == 0) L.Wo = retrace_weight(B, L, 1, ov, L.stored, K);   // (own weight + StoredR) / K
  __syncthreads();
  // retrace: orientations about the old first bead, trial 0 its own
  if (chain > 0) {
    for (int t = tid; t < O * chain; t += kBT) {
      const int o = t / chain, a = t - o * chain;
      BSite s;
      if (o == 0) {
        const int j = L.start + 1 + a;
        s.x = B.g[j]; s.y = B.g[cap + j]; s.z = B.g[2 * cap + j];
      } else {
        const double* t0 = L.tmpl[comp][0];
        const double* t1 = L.tmpl[comp][1 + a];
        double v[3] = {t1[0] - t0[0], t1[1] - t0[1], t1[2] - t0[2]};
        const int d = kBDrawTrial + 3 * K + 3 * O + 3 * o;
        quat_rotate(v, L.ud[d], L.ud[d + 1], L.ud[d + 2]);
        s.x = B.g[L.start] + v[0]; s.y = B.g[cap + L.start] + v[1]; s.z = B.g[2 * cap + L.start] + v[2];
      }
      s.q = sp.q[1 + a];
      s.type = sp.type[1 + a];
      s._p = 0;
      L.site[t] = s;
    }
    __syncthreads();
    ov = trial_energies<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, O, chain, ex_lo, ex_hi);
    if (tid == 0) L.Wo *= retrace_weight(B, L, O, ov, 0.0, O);
    __syncthreads();
  }
  // Ewald change of moving the molecule: old sites, then new ones
  if (L.charged) {
    if (tid < 2 * M) {
      const int i = tid < M ? tid : tid - M;
      BSite s;
      if (tid < M) {
        const int j = L.start + i;
        s.x = B.g[j]; s.y = B.g[cap + j]; s.z = B.g[2 * cap + j];
      } else {
        s.x = L.molp[i][0]; s.y = L.molp[i][1]; s.z = L.molp[i][2];
      }
      s.q = sp.q[i];
      s.type = sp.type[i];
      s._p = 0;
      L.site[tid] = s;
    }
    __syncthreads();
    ewald_pass(B, S, L, eik, M, M);
  }
  if (tid == 0) {
    L.en[kHGV] -= L.eo[0]; L.en[kHGR] -= L.eo[1];
    L.en[kGGV] -= L.eo[2]; L.en[kGGR] -= L.eo[3];
    double W = L.W;
    if (L.charged) {
      L.en[kGGE] = L.same;
      L.en[kHGE] = 2.0 * L.cross;
      W *= exp(-B.beta * (L.same + 2.0 * L.cross));
    }
    L.W = W;
    const double pacc = W / L.Wo;
    L.pacc = pacc;
    const int acc = (W > 1e-150 && L.ud[4] < pacc) ? 1 : 0;
    L.accept = acc;
    if (acc) {
      S.cnt[comp][10]++;
      add_energy(S.running, L.en);
      if (L.charged) S.sk_cur ^= 1;
    }
  }
  __syncthreads();
  if (L.accept && tid < M) {
    const int j = L.start + tid;
    B.g[j] = L.molp[tid][0]; B.g[cap + j] = L.molp[tid][1]; B.g[2 * cap + j] = L.molp[tid][2];
    mirror_pos(B, L, gml, comp, j, L.molp[tid][0], L.molp[tid][1], L.molp[tid][2]);
  }
}

// Identity swap (IdentitySwapMove, mc_swap_moves.h:199-426): the OLD
// molecule L.mol of species L.comp is replaced by one of species L.nc grown on
// its first bead (one trial whose weight is not divided by K, then O
// orientations of the NEW template about it), energies without the OLD
// molecule; the OLD molecule is retraced with its own first bead and O
// orientations (trial 0 its own).  W_new carries the Ewald change of the OLD
// sites out and the NEW sites in and the tail change; Pacc = pref_ins(NEW)
// pref_del(OLD) (W_new / W_id,NEW) / (W_old / W_id,OLD).  On acceptance the
// OLD molecule is deleted (the last one into its hole) and the NEW one
// appended, or, for the same species, moved in place.  Draws: orientation o
// at kBDrawTrial + 3K + 3o, retrace orientation o at kBDrawTrial + 3K + 3O + 3o.
template <bool MIXED, bool CUBIC>
__device__ void identity_swap(const BBox& B, BState& S, BMove& L, const double* ffl, const float* thr,
                              const double2* ct, const double* fwl, double* gml, int* loc, double2* eik)
{
  const int tid = threadIdx.x;
  const int oc = L.comp, nc = L.nc, cap = B.guest_cap;
  const BSpecies& so = B.sp[oc];
  const BSpecies& sn = B.sp[nc];
  const int Mo = so.molsize, Mn = sn.molsize, K = B.K, O = B.O;
  const int ex_lo = L.start, ex_hi = L.start + Mo;
  // NEW first bead at the OLD molecule's first bead
  if (tid == 0) {
    BSite s;
    s.x = B.g[L.start]; s.y = B.g[cap + L.start]; s.z = B.g[2 * cap + L.start];
    s.q = sn.q[0];
    s.type = sn.type[0];
    s._p = 0;
    L.site[0] = s;
    for (int t = 0; t < 4; ++t) L.eo[t] = 0.0;
  }
  __syncthreads();
  uint32_t ov = trial_energies<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, 1, 1, ex_lo, ex_hi);
  if (tid < 64) {
    double W = 1.0;
    const int sel = select_wave0(B, L, 1, ov, 0.5, false, false, &W);
    if (tid == 0) {
      if (sel < 0) {
        L.fail = 1;
      } else {
        L.W = W;
        L.en[kHGV] = L.t4[0][0]; L.en[kHGR] = L.t4[0][1];
        L.en[kGGV] = L.t4[0][2]; L.en[kGGR] = L.t4[0][3];
        L.molp[0][0] = L.site[0].x; L.molp[0][1] = L.site[0].y; L.molp[0][2] = L.site[0].z;
      }
    }
  }
  __syncthreads();
  if (L.fail) return;
  // NEW chain: O orientations of the NEW template
  if (Mn > 1) {
    const int chain = Mn - 1;
    for (int t = tid; t < O * chain; t += kBT) {
      const int o = t / chain, a = t - o * chain;
      const double* t0 = L.tmpl[nc][0];
      const double* t1 = L.tmpl[nc][1 + a];
      double v[3] = {t1[0] - t0[0], t1[1] - t0[1], t1[2] - t0[2]};
      const int d = kBDrawTrial + 3 * K + 3 * o;
      quat_rotate(v, L.ud[d], L.ud[d + 1], L.ud[d + 2]);
      BSite s;
      s.x = L.molp[0][0] + v[0]; s.y = L.molp[0][1] + v[1]; s.z = L.molp[0][2] + v[2];
      s.q = sn.q[1 + a];
      s.type = sn.type[1 + a];
      s._p = 0;
      L.site[t] = s;
    }
    __syncthreads();
    ov = trial_energies<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, O, chain, ex_lo, ex_hi);
    if (tid < 64) {
      double W = L.W;
      const int sel = select_wave0(B, L, O, ov, L.ud[6], false, true, &W);
      if (tid == 0) {
        if (sel < 0) {
          L.fail = 1;
        } else {
          L.W = W;
          L.en[kHGV] += L.t4[sel][0]; L.en[kHGR] += L.t4[sel][1];
          L.en[kGGV] += L.t4[sel][2]; L.en[kGGR] += L.t4[sel][3];
          for (int a = 0; a < chain; ++a) {
            const BSite& s = L.site[sel * chain + a];
            L.molp[1 + a][0] = s.x; L.molp[1 + a][1] = s.y; L.molp[1 + a][2] = s.z;
          }
        }
      }
    }
    __syncthreads();
    if (L.fail) return;
  }
  // retrace of the OLD molecule: its own first bead
  if (tid == 0) {
    BSite s;
    s.x = B.g[L.start]; s.y = B.g[cap + L.start]; s.z = B.g[2 * cap + L.start];
    s.q = so.q[0];
    s.type = so.type[0];
    s._p = 0;
    L.site[0] = s;
  }
  __syncthreads();
  ov = trial_energies<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, 1, 1, ex_lo, ex_hi);
  if (tid == 0) L.Wo = retrace_weight(B, L, 1, ov, 0.0, 1);
  __syncthreads();
  if (Mo > 1) {
    const int chain = Mo - 1;
    for (int t = tid; t < O * chain; t += kBT) {
      const int o = t / chain, a = t - o * chain;
      BSite s;
      if (o == 0) {
        const int j = L.start + 1 + a;
        s.x = B.g[j]; s.y = B.g[cap + j]; s.z = B.g[2 * cap + j];
      } else {
        const double* t0 = L.tmpl[oc][0];
        const double* t1 = L.tmpl[oc][1 + a];
        double v[3] = {t1[0] - t0[0], t1[1] - t0[1], t1[2] - t0[2]};
        const int d = kBDrawTrial + 3 * K + 3 * O + 3 * o;
        quat_rotate(v, L.ud[d], L.ud[d + 1], L.ud[d + 2]);
        s.x = B.g[L.start] + v[0]; s.y = B.g[cap + L.start] + v[1]; s.z = B.g[2 * cap + L.start] + v[2];
      }
      s.q = so.q[1 + a];
      s.type = so.type[1 + a];
      s._p = 0;
      L.site[t] = s;
    }
    __syncthreads();
    ov = trial_energies<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, O, chain, ex_lo, ex_hi);
    if (tid == 0) L.Wo *= retrace_weight(B, L, O, ov, 0.0, O);
    __syncthreads();
  }
  // Ewald: the OLD sites out (charged species only), then the NEW sites in
  const int n_old = L.qo ? Mo : 0, n_new = L.qn ? Mn : 0;
  if (n_old + n_new > 0) {
    if (tid < n_old + n_new) {
      BSite s;
      if (tid < n_old) {
        const int j = L.start + tid;
        s.x = B.g[j]; s.y = B.g[cap + j]; s.z = B.g[2 * cap + j];
        s.q = so.q[tid];
        s.type = so.type[tid];
      } else {
        const int i = tid - n_old;
        s.x = L.molp[i][0]; s.y = L.molp[i][1]; s.z = L.molp[i][2];
        s.q = sn.q[i];
        s.type = sn.type[i];
      }
      s._p = 0;
      L.site[tid] = s;
    }
    __syncthreads();
    ewald_pass(B, S, L, eik, n_old, n_new);
  }
  if (tid == 0) {
    L.en[kHGV] -= L.eo[0]; L.en[kHGR] -= L.eo[1];
    L.en[kGGV] -= L.eo[2]; L.en[kGGR] -= L.eo[3];
    double W = L.W;
    if (n_old + n_new > 0) {
      double same = L.same;
      if (L.qn) same -= sn.excl * 1.0;
      if (L.qo) same -= so.excl * -1.0;
      L.en[kGGE] = same;
      L.en[kHGE] = 2.0 * L.cross;
      W *= exp(-B.beta * (same + 2.0 * L.cross));
    }
    if (B.has_tail) L.en[kTail] = tail_identity(B, S, nc, oc);
    W *= exp(-B.beta * L.en[kTail]);
    L.W = W;
    const double V = B.volume;
    double pref = B.beta * sn.mol_fraction * B.pressure * sn.fugacity * V / (1.0 + (double)S.n_mol[nc]);
    pref *= (double)S.n_mol[oc] / (B.beta * so.mol_fraction * B.pressure * so.fugacity * V);
    const double pacc = pref * (W / sn.ideal_rosen) / (L.Wo / so.ideal_rosen);
    L.pacc = pacc;
    int acc = (L.ud[4] < pacc) ? 1 : 0;
    if (acc && nc != oc && S.n_mol[nc] + 1 > sn.capacity) {
      S.error = GRASPA_ERR_CAPACITY;
      acc = 0;
    }
    L.accept = acc;
    if (acc) {
      S.cnt[oc][GRASPA_CNT_IDSW_ACC + nc - 1]++;
      if (nc != oc)
        for (int t = 0; t < GRASPA_MAX_PSEUDO; ++t) S.n_pseudo[t] += sn.pseudo[t] - so.pseudo[t];
      add_energy(S.running, L.en);
      if (n_old + n_new > 0) S.sk_cur ^= 1;
    }
  }
  __syncthreads();
  if (!L.accept) return;
  if (nc != oc) {
    // the last OLD molecule into the hole (its MolIDs stay), the NEW one appended
    const int last = so.base + (S.n_mol[oc] - 1) * Mo;
    if (L.start != last && tid < Mo) {
      const int j = L.start + tid, k = last + tid;
      const double x = B.g[k], y = B.g[cap + k], z = B.g[2 * cap + k], q = B.g[3 * cap + k];
      const int t = B.g_type[k];
      B.g[j] = x; B.g[cap + j] = y; B.g[2 * cap + j] = z;
      B.g[3 * cap + j] = q;
      B.g_type[j] = t;
      mirror_put(B, L, gml, oc, j, x, y, z, q, t);
    }
    const int nn = S.n_mol[nc];
    if (tid >= 64 && tid < 64 + Mn) {
      const int a = tid - 64, j = sn.base + nn * Mn + a;
      B.g[j] = L.molp[a][0]; B.g[cap + j] = L.molp[a][1]; B.g[2 * cap + j] = L.molp[a][2];
      B.g[3 * cap + j] = sn.q[a];
      B.g_type[j] = sn.type[a];
      B.g_molid[j] = nn;
      B.g_comp[j] = nc;
      mirror_put(B, L, gml, nc, j, L.molp[a][0], L.molp[a][1], L.molp[a][2], sn.q[a], sn.type[a]);
    }
    __syncthreads();
    if (tid == 0) {
      S.n_mol[oc] -= 1;
      S.n_mol[nc] += 1;
    }
  } else if (tid < Mo) {
    const int j = L.start + tid;
    B.g[j] = L.molp[tid][0]; B.g[cap + j] = L.molp[tid][1]; B.g[2 * cap + j] = L.molp[tid][2];
    mirror_pos(B, L, gml, oc, j, L.molp[tid][0], L.molp[tid][1], L.molp[tid][2]);
  }
}

// Update_Max_Translation / Update_Max_Rotation (mc_utilities.h:612-665)
__device__ inline void update_max(double* v, int64_t& acc, int64_t& tot, double hi)
{
  const double ratio = (double)acc / (double)tot;
  for (int d = 0; d < 3; ++d) {
    v[d] *= (ratio > 0.5) ? 1.05 : 0.95;
    if (v[d] < 0.01) v[d] = 0.01;
    if (v[d] > hi) v[d] = hi;
  }
  acc = 0;
  tot = 0;
}

template <bool MIXED, bool CUBIC>
__global__ __launch_bounds__(kBT) void batch_mc_kernel(const BBox* __restrict__ boxes, BState* __restrict__ states,
                                                       const int max_moves, const int64_t* __restrict__ budgets,
                                                       const int dyn_bytes)
{
  extern __shared__ double4 dyn4[];
  __shared__ BState S;
  __shared__ BMove L;
  __shared__ BBox B;
  __shared__ int cyc_block;          // block of a production cycle that just ended, else -1
  if (threadIdx.x == 0) cyc_block = -1;
  double* dyn = reinterpret_cast<double*>(dyn4);
  const int tid = threadIdx.x;
  {
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&states[blockIdx.x]);
    uint64_t* dst = reinterpret_cast<uint64_t*>(&S);
    for (int i = tid; i < (int)(sizeof(BState) / 8); i += kBT) dst[i] = src[i];
    const uint64_t* bsrc = reinterpret_cast<const uint64_t*>(&boxes[blockIdx.x]);
    uint64_t* bdst = reinterpret_cast<uint64_t*>(&B);
    for (int i = tid; i < (int)(sizeof(BBox) / 8); i += kBT) bdst[i] = bsrc[i];
  }
  __syncthreads();
  const int T = B.bc.ntypes;
  const int nff = 4 * T * T;
  const int nff4 = (nff + 3) & ~3;
  const int nthr = ((T * T + 7) & ~7) / 2;     // screen thresholds, in doubles
  double* ffl = dyn;
  float* thr = reinterpret_cast<float*>(dyn + nff4);
  double2* ct = reinterpret_cast<double2*>(dyn + nff4 + nthr);
  const int nct = MIXED ? ((coul_lds_doubles(B.coul_n) + 3) & ~3) : 0;
  double2* eik = reinterpret_cast<double2*>(dyn + nff4 + nthr + nct);
  // the framework goes to LDS for the launch when it fits after the tables
  const int fw_off = nff4 + nthr + nct + 2 * (2 * kBMaxMol * 3 * B.kstride);
  const int nf = B.n_fw;
  const bool fw_in_lds = (size_t)fw_off * 8 + (size_t)nf * 36 <= (size_t)dyn_bytes;
  double* fwl = fw_in_lds ? dyn + fw_off : nullptr;
  if (fw_in_lds) {
    for (int i = tid; i < 4 * nf; i += kBT) fwl[i] = B.fw[i];
    int* ftl = reinterpret_cast<int*>(fwl + 4 * nf);
    for (int i = tid; i < nf; i += kBT) ftl[i] = B.fw_type[i];
  }
  // the guest atoms go to an LDS mirror in what the framework leaves: every
  // species gets an equal share, at most its capacity of molecules; the pair
  // passes read a species from the mirror while its atoms fit there
  // then the pair passes' neighbour-list segments (cubic cells), then the
  // guest mirror
  const int l_off = fw_in_lds ? fw_off + 4 * nf + (nf + 1) / 2 : fw_off;
  constexpr int kLocDoubles = kBW * (kLocFw + kLocG) / 2;
  int* loc = (CUBIC && (size_t)(l_off + kLocDoubles) * 8 <= (size_t)dyn_bytes) ? reinterpret_cast<int*>(dyn + l_off)
                                                                               : nullptr;
  const int g_off = loc ? l_off + kLocDoubles : l_off;
  const int g_room = dyn_bytes / 8 > g_off ? (int)((dyn_bytes / 8 - g_off) * 8 / 36) : 0;
  if (tid == 0) {
    const int ns = B.n_comp - 1, share = ns > 0 ? g_room / ns : 0;
    int n = 0;
    for (int c = 0; c < kMaxComps; ++c) {
      int cp = 0;
      if (c >= 1 && c < B.n_comp) {
        const int M = B.sp[c].molsize;
        cp = min(B.sp[c].capacity, share / M) * M;
      }
      L.gl_off[c] = n;
      L.gl_cap[c] = cp;
      n += cp;
    }
    L.gl_n = n;
  }
  __syncthreads();
  double* gml = L.gl_n > 0 ? dyn + g_off : nullptr;
  if (gml) {
    const int gn = L.gl_n;
    int* gtl = reinterpret_cast<int*>(gml + 4 * gn);
    for (int i = tid; i < gn; i += kBT) {
      int c = 1;
      while (c + 1 < B.n_comp && i >= L.gl_off[c + 1]) ++c;
      const int k = i - L.gl_off[c];
      if (k < S.n_mol[c] * B.sp[c].molsize) {
        const int j = B.sp[c].base + k, cap = B.guest_cap;
        gml[i] = B.g[j]; gml[gn + i] = B.g[cap + j]; gml[2 * gn + i] = B.g[2 * cap + j];
        gml[3 * gn + i] = B.g[3 * cap + j];
        gtl[i] = B.g_type[j];
      }
    }
  }
  for (int i = tid; i < nff; i += kBT) ffl[i] = B.ff64[i];
  for (int i = tid; i < T * T; i += kBT) {
    const int st = i / T, at = i - st * T;
    thr[i] = screen_threshold(B.ff64 + 4 * (at * T + st), B.bc.overlap, B.bc.cut_vdw2);
  }
  if (MIXED)
    for (int i = tid; i < B.coul_n / 2; i += kBT)
      ct[(i / (kCoulTabStride / 2)) * kCoulLdsRow + i % (kCoulTabStride / 2)] = reinterpret_cast<const double2*>(B.coul)[i];
  __syncthreads();
  const int nd = kBDrawTrial + 3 * B.K + 6 * B.O;
  const int64_t limit = budgets ? budgets[blockIdx.x] : (int64_t)max_moves;
  const uint64_t t_launch = wall_clock64();
  for (int64_t mv = 0; mv < limit; ++mv) {
    if (S.cycle >= S.target || S.error) break;
    // this move's uniforms, one Philox block (two uniforms) per thread
    if (2 * tid < nd) {
      uint32_t c[4] = {(uint32_t)S.move, (uint32_t)((uint64_t)S.move >> 32), (uint32_t)tid, B.stream};
      philox4x32_10(c, (uint32_t)B.seed, (uint32_t)(B.seed >> 32));
      L.ud[2 * tid] = ((double)((((uint64_t)c[0] << 32) | c[1]) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
      L.ud[2 * tid + 1] = ((double)((((uint64_t)c[2] << 32) | c[3]) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
    if (tid >= 256 && tid < 256 + (kMaxComps - 1) * kBMaxMol * 3) {
      // slot 0 of every species (the orientation template, mc_widom.h:256)
      const int t = tid - 256, c = 1 + t / (kBMaxMol * 3), a = (t / 3) % kBMaxMol, d = t % 3;
      if (c < B.n_comp && a < B.sp[c].molsize) L.tmpl[c][a][d] = B.g[d * B.guest_cap + B.sp[c].base + a];
    }
    if (tid == 0 && S.step == 0) {
      // Determine_Number_Of_Steps (axpy.cu:580-591)
      int64_t steps = 20;
      if (steps < S.total_mol) steps = S.total_mol;
      if (B.use_max_step && steps > B.max_step) steps = B.max_step;
      S.steps = (int)steps;
      S.cycle_t0 = wall_clock64();
    }
    __syncthreads();
    if (tid == 0) {
      const double* u = L.ud;
      int vi = (int)(u[0] * (double)B.n_valid);
      if (vi >= B.n_valid) vi = B.n_valid - 1;
      const int comp = B.valid[vi];
      const int n = S.n_mol[comp];
      const int mol = (int)(u[1] * (double)n);
      const double r = u[2];
      const BSpecies& sp = B.sp[comp];
      int tm = kMvNone, work = 0;
      if (r < sp.p_trans) { tm = kMvTrans; work = n > 0; }
      else if (r < sp.p_rot) { tm = kMvRot; work = n > 0; }
      else if (r < sp.p_widom) { tm = kMvWidom; work = 1; }
      else if (r < sp.p_reins) { tm = kMvReins; work = n > 0; }
      else if (r < sp.p_ident) {           // attempted while the box holds an adsorbate
        tm = kMvIdent;
        for (int c = 1; c < B.n_comp; ++c) work |= S.n_mol[c] > 0 ? 1 : 0;
      }
      else if (r < sp.p_swap) {
        if (u[3] < 0.5) { tm = kMvIns; work = 1; }
        else { tm = kMvDel; work = n > 0; }
      }
      L.tmove = tm;
      L.type = work ? tm : kMvNone;
      L.comp = comp;
      L.mol = mol;
      L.M = sp.molsize;
      L.start = sp.base + mol * sp.molsize;
      L.charged = sp.charged;
      for (int i = 0; i < 10; ++i) L.en[i] = 0.0;
      L.W = 0.0; L.pacc = 0.0; L.same = 0.0; L.cross = 0.0;
      L.fail = 0; L.accept = 0; L.overlap = 0;
      L.nc = 0; L.qo = 0; L.qn = 0;
      if (work && tm == kMvIdent) {
        // the species pair of IdentitySwapMove's rejection loop
        // (mc_swap_moves.h:216-235): old uniform over the species holding
        // molecules, new uniform over all; then the old molecule
        int held[kMaxComps], nh = 0;
        for (int c = 1; c < B.n_comp; ++c)
          if (S.n_mol[c] > 0) held[nh++] = c;
        int i = (int)(u[kBDrawIdOld] * (double)nh);
        if (i >= nh) i = nh - 1;
        const int oc = held[i];
        int nc = 1 + (int)(u[kBDrawIdNew] * (double)(B.n_comp - 1));
        if (nc >= B.n_comp) nc = B.n_comp - 1;
        const BSpecies& so = B.sp[oc];
        L.comp = oc;
        L.nc = nc;
        L.mol = (int)(u[kBDrawIdMol] * (double)S.n_mol[oc]);
        L.M = so.molsize;
        L.start = so.base + L.mol * so.molsize;
        L.qo = so.charged;
        L.qn = B.sp[nc].charged;
        L.charged = L.qo | L.qn;
        S.cnt[oc][GRASPA_CNT_IDSW_TOT + nc - 1]++;
      } else if (work) {
        if (tm == kMvTrans) { S.cnt[comp][0]++; S.upd_tot[comp][0]++; }
        else if (tm == kMvRot) { S.cnt[comp][2]++; S.upd_tot[comp][1]++; }
        else if (tm == kMvIns) S.cnt[comp][4]++;
        else if (tm == kMvDel) S.cnt[comp][6]++;
        else if (tm == kMvWidom) S.cnt[comp][8]++;
        else if (tm == kMvReins) S.cnt[comp][9]++;
      }
    }
    __syncthreads();
    const int type = L.type;
    if (type == kMvTrans || type == kMvRot) single_body<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, eik);
    else if (type == kMvReins) reinsertion<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, eik);
    else if (type == kMvIdent) identity_swap<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, eik);
    else if (type != kMvNone) cbmc<MIXED, CUBIC>(B, S, L, ffl, thr, ct, fwl, gml, loc, eik);
    __syncthreads();
    if (tid == 0) {
      const int comp = L.comp;
      if (L.fail && L.type == kMvWidom && S.production) {
        S.widom_n[comp] += 1.0;
        B.blk->widom_n[block_of(S, B.n_block)][comp] += 1.0;
      }
      if (B.trace && S.trace_n < B.trace_cap) {
        graspa_trace& tr = B.trace[(size_t)blockIdx.x * B.trace_cap + S.trace_n];
        tr.cycle = (int)S.cycle; tr.step = S.step; tr.move = L.tmove; tr.comp = comp; tr.mol = L.mol;
        tr.accepted = L.accept; tr.overlap = L.overlap; tr._pad = 0;
        tr.pacc = L.pacc;
        tr.rosen = L.fail ? 0.0 : L.W;
        const bool keep = (L.accept || L.type == kMvWidom) && !L.fail;
        double* de = &tr.de.hh_vdw;
        for (int i = 0; i < 10; ++i) de[i] = keep ? L.en[i] : 0.0;
      }
      S.trace_n++;
      S.move++;
      S.step++;
      if (S.step >= S.steps) {
        // end of cycle: loadings, max-step update (axpy.cu:378-400), latency
        if (S.production) {
          for (int c = 1; c < B.n_comp; ++c) {
            const double nm = (double)S.n_mol[c];
            S.load_sum[c] += nm;
            S.load_sum2[c] += nm * nm;
          }
          S.load_n++;
          cyc_block = block_of(S, B.n_block);     // block sums by all threads after the barrier
        }
        if (S.cycle % 500 == 0) {
          for (int c = 1; c < B.n_comp; ++c) {
            if (S.upd_tot[c][0] > 0) update_max(S.max_trans[c], S.upd_acc[c][0], S.upd_tot[c][0], 5.0);
            if (S.upd_tot[c][1] > 0) update_max(S.max_rot[c], S.upd_acc[c][1], S.upd_tot[c][1], 3.14);
          }
        }
        if (B.cyc) B.cyc[S.cycle % B.cyc_cap] = (float)(wall_clock64() - S.cycle_t0);
        S.cycle++;
        S.step = 0;
      }
    }
    __syncthreads();
    if (cyc_block >= 0) {             // set by thread 0 before the barrier: the same for all threads
      block_sums(B, S, cyc_block, tid);
      __syncthreads();
      if (tid == 0) cyc_block = -1;
    }
  }
  if (threadIdx.x == 0) S.busy += wall_clock64() - t_launch;
  __syncthreads();
  {
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&S);
    uint64_t* dst = reinterpret_cast<uint64_t*>(&states[blockIdx.x]);
    for (int i = tid; i < (int)(sizeof(BState) / 8); i += kBT) dst[i] = src[i];
  }
}

}  // namespace graspa

using namespace graspa;

struct graspa_batch {
  int device = 0;
  int precision = GRASPA_FP64;
  int cubic = 1;
  hipStream_t stream = nullptr;
  int n = 0;
  std::vector<graspa_box*> boxes;   // slot -> box (nullptr: empty slot)
  std::vector<BBox> hb;
  std::vector<BState> hs;
  BBox* d_box = nullptr;
  BState* d_state = nullptr;
  std::vector<double2*> skc;        // per slot: framework S(k), adsorbate S(k) x 2 (active k-vectors)
  double* d_tail = nullptr;         // kMaxTypes^2 per slot
  int* d_use_tail = nullptr;
  graspa_trace* d_trace = nullptr;
  int trace_cap = 0;
  float* d_cyc = nullptr;
  int cyc_cap = 0;
  graspa_block_stats* d_blk = nullptr;   // per slot
  size_t lds = 0;
  int max_lds = 65536;
  double tick_hz = 1.0e8;
  double kernel_ms = 0.0;
  int64_t launches = 0;
  // the launch in flight (graspa_batch_launch_async): its events
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool in_flight = false;
  // time slices (graspa_batch_set_time_slice): per-box move budgets of a
  // launch from each box's measured time per move
  double slice_us = 0.0;
  std::vector<int64_t> hbud;
  int64_t* d_bud = nullptr;
  std::vector<double> us_per_move;   // running estimate per slot (0: none yet)
  std::vector<int64_t> seen_moves;
  std::vector<uint64_t> seen_busy;
};

namespace {

int finish_launch(graspa_batch* bt);
int copy_states_back(graspa_batch* bt);

int upload_batch(graspa_batch* bt)
{
  if (int rc = finish_launch(bt)) return rc;
  if (hip_check(hipMemcpyAsync(bt->d_box, bt->hb.data(), bt->n * sizeof(BBox), hipMemcpyHostToDevice, bt->stream),
                "batch upload"))
    return GRASPA_ERR_HIP;
  if (hip_check(hipMemcpyAsync(bt->d_state, bt->hs.data(), bt->n * sizeof(BState), hipMemcpyHostToDevice, bt->stream),
                "batch upload"))
    return GRASPA_ERR_HIP;
  return hip_check(hipStreamSynchronize(bt->stream), "batch upload");
}

int download_states(graspa_batch* bt)
{
  if (int rc = finish_launch(bt)) return rc;
  if (int rc = copy_states_back(bt)) return rc;
  // time per move of each box over the moves since the last look
  for (int i = 0; i < bt->n; ++i) {
    const BState& st = bt->hs[i];
    const int64_t dm = st.move - bt->seen_moves[i];
    if (bt->boxes[i] && dm > 0) {
      const double us = (double)(st.busy - bt->seen_busy[i]) / bt->tick_hz * 1e6 / (double)dm;
      bt->us_per_move[i] = bt->us_per_move[i] > 0.0 ? 0.5 * (bt->us_per_move[i] + us) : us;
    }
    bt->seen_moves[i] = st.move;
    bt->seen_busy[i] = st.busy;
  }
  return GRASPA_OK;
}

int copy_states_back(graspa_batch* bt)
{
  if (hip_check(hipMemcpyAsync(bt->hs.data(), bt->d_state, bt->n * sizeof(BState), hipMemcpyDeviceToHost, bt->stream),
                "batch state copy"))
    return GRASPA_ERR_HIP;
  return hip_check(hipStreamSynchronize(bt->stream), "batch state copy");
}

void free_batch(graspa_batch* bt)
{
  if (!bt) return;
  hipSetDevice(bt->device);
  if (bt->stream) hipStreamSynchronize(bt->stream);
  if (bt->ev0) hipEventDestroy(bt->ev0);
  hipFree(bt->d_bud);
  if (bt->ev1) hipEventDestroy(bt->ev1);
  hipFree(bt->d_box); hipFree(bt->d_state); hipFree(bt->d_tail); hipFree(bt->d_use_tail);
  hipFree(bt->d_trace); hipFree(bt->d_cyc); hipFree(bt->d_blk);
  for (double2* p : bt->skc) hipFree(p);
  if (bt->stream) hipStreamDestroy(bt->stream);
  delete bt;
}

// Dynamic LDS of one box: force field, real-space table (fast form), e^{ik.r}
// tables of up to 2 x 8 sites.
size_t box_lds(const graspa_box* b, int precision)
{
  const int T = b->bc.ntypes;
  const int ks = 1 + std::max(b->bc.kmax[0], std::max(b->bc.kmax[1], b->bc.kmax[2]));
  const size_t nff4 = ((size_t)4 * T * T + 3) & ~(size_t)3;
  const size_t nthr = (((size_t)T * T + 7) & ~(size_t)7) / 2;
  const size_t nct = precision == GRASPA_MIXED ? (((size_t)coul_lds_doubles(b->coul_n) + 3) & ~(size_t)3) : 0;
  return (nff4 + nthr + nct) * sizeof(double) + (size_t)2 * kBMaxMol * 3 * ks * sizeof(double2);
}

// Validates one box and fills slot i's descriptor and initial state (host
// side); the caller uploads.
int fill_slot(graspa_batch* bt, int i, const graspa_batch_box& in)
{
  graspa_box* b = in.box;
  const graspa_mc_params* p = in.params;
  if (!b || !p) return fail(GRASPA_ERR_ARG, "null box or params");
  if (in.n_block < 0 || in.n_block > GRASPA_NBLOCK) return fail(GRASPA_ERR_ARG, "n_block must be in 0..16");
  if (b->device != bt->device || b->precision != bt->precision)
    return fail(GRASPA_ERR_ARG, "all boxes of a batch must share the device and the precision form");
  const int ncomp = (int)b->comps.size();
  if (p->n_species != ncomp - 1 || ncomp < 2)
    return fail(GRASPA_ERR_ARG, "n_species must match the box's adsorbate components");
  if (p->n_trial_pos < 1 || p->n_trial_pos > kBMaxTrial || p->n_trial_orient < 1 || p->n_trial_orient > kBMaxTrial)
    return fail(GRASPA_ERR_ARG, "trial counts must be in 1..32");
  const int T = b->bc.ntypes;
  if (T > kMaxTypes) return fail(GRASPA_ERR_ARG, "too many atom types");
  if (b->bc.use_charges && b->nvec <= 0) return fail(GRASPA_ERR_STATE, "charged box without Ewald tables");
  if (p->has_tail && (!p->tail || !p->use_tail)) return fail(GRASPA_ERR_ARG, "tail tables required when has_tail");
  const size_t need = box_lds(b, bt->precision);
  if (need + sizeof(BState) + sizeof(BMove) + sizeof(BBox) > (size_t)bt->max_lds)
    return fail(GRASPA_ERR_UNSUPPORTED, "box workspace exceeds LDS (force field, real-space table, k range)");
  BBox h;
  std::memset(&h, 0, sizeof(h));
  h.bc = b->bc;
  h.volume = b->volume;
  // the framework copy sorted by the Morton code of its screen cells: 64
  // consecutive atoms are close in space, so in a pair pass the lanes of a
  // wave are mostly all inside or all beyond the cutoff of a site and the
  // pair terms are skipped for whole waves of far atoms
  h.fw = b->d_fws ? b->d_fws : b->d_fw;
  h.fw_type = b->d_fws ? b->d_fws_type : b->d_fw_type;
  h.n_fw = b->n_fw;
  h.guest_cap = b->guest_cap;
  h.g = b->d_g; h.g_type = b->d_g_type; h.g_molid = b->d_g_molid; h.g_comp = b->d_g_comp;
  h.ff64 = b->d_ff64;
  h.coul = b->d_coul; h.coul_n = b->coul_n;
  h.kstride = 1 + std::max(b->bc.kmax[0], std::max(b->bc.kmax[1], b->bc.kmax[2]));
  h.kact = b->d_kact; h.ktemp_act = b->d_ktemp_act; h.n_active = b->bc.use_charges ? b->n_active : 0;
  hipFree(bt->skc[i]);
  bt->skc[i] = nullptr;
  if (h.n_active > 0) {
    // structure factors of the active k-vectors in the batch's own compact
    // arrays (one load per k-vector, no index gather); written back by sync
    const int na = h.n_active, nv = b->nvec;
    std::vector<int4> ka(na);
    std::vector<double2> fw(nv), ads(nv), c(3 * (size_t)na);
    if (hip_check(hipMemcpy(ka.data(), b->d_kact, na * sizeof(int4), hipMemcpyDeviceToHost), "k copy") ||
        hip_check(hipMemcpy(fw.data(), b->d_sk_fw, nv * sizeof(double2), hipMemcpyDeviceToHost), "sk copy") ||
        hip_check(hipMemcpy(ads.data(), b->d_sk_ads, nv * sizeof(double2), hipMemcpyDeviceToHost), "sk copy"))
      return GRASPA_ERR_HIP;
    for (int j = 0; j < na; ++j) {
      c[j] = fw[ka[j].w];
      c[na + j] = ads[ka[j].w];
      c[2 * (size_t)na + j] = ads[ka[j].w];
    }
    if (hip_check(hipMalloc(&bt->skc[i], 3 * (size_t)na * sizeof(double2)), "hipMalloc batch")) return GRASPA_ERR_HIP;
    if (hip_check(hipMemcpy(bt->skc[i], c.data(), 3 * (size_t)na * sizeof(double2), hipMemcpyHostToDevice), "sk copy"))
      return GRASPA_ERR_HIP;
    h.skfw = bt->skc[i];
    h.skc[0] = bt->skc[i] + na;
    h.skc[1] = bt->skc[i] + 2 * (size_t)na;
  }
  if (b->use_cells) {
    h.ctab = b->d_ctab;
    h.cell_cap = b->cell_cap;
    for (int d = 0; d < 3; ++d) h.cell_n[d] = b->cell_n[d];
  }
  h.n_comp = ncomp;
  for (int c = 1; c < ncomp; ++c) {
    const graspa_species& sp = p->species[c - 1];
    const Component& cc = b->comps[c];
    if (cc.molsize > kBMaxMol) return fail(GRASPA_ERR_ARG, "molecules of at most 8 atoms");
    if ((cc.molsize - 1) * p->n_trial_orient > kBMaxSites)
      return fail(GRASPA_ERR_ARG, "n_trial_orient * (molsize - 1) must be <= 64");
    BSpecies& d = h.sp[c];
    d.p_trans = sp.p_translation; d.p_rot = sp.p_rotation; d.p_widom = sp.p_widom; d.p_swap = sp.p_swap;
    d.p_reins = sp.p_reinsertion;
    d.p_ident = sp.p_identity;
    d.mol_fraction = sp.mol_fraction; d.fugacity = sp.fugacity_coeff; d.ideal_rosen = sp.ideal_rosen;
    d.excl = sp.excl_intra + sp.excl_self;
    d.base = cc.base; d.molsize = cc.molsize; d.capacity = cc.capacity;
    double qabs = 0.0;
    for (int a = 0; a < cc.molsize; ++a) qabs += std::fabs(cc.q[a]);
    d.charged = (b->bc.use_charges && qabs > 1e-6) ? 1 : 0;   // hasPartialCharge (read_data.cpp:2160-2164)
    for (int a = 0; a < cc.molsize; ++a) {
      d.q[a] = cc.q[a];
      d.type[a] = cc.type[a];
    }
    for (int t = 0; t < GRASPA_MAX_PSEUDO; ++t) d.pseudo[t] = (int)sp.pseudo_per_mol[t];
    d.chain = cc.d_chain;
    if (cc.d_chain) {
      // flexible chains grow by insertion / deletion / Widom; reinsertion and
      // identity swaps of a flexible molecule are not implemented
      if (sp.p_reinsertion > sp.p_widom || sp.p_identity > std::max(sp.p_reinsertion, sp.p_widom))
        return fail(GRASPA_ERR_UNSUPPORTED, "flexible chains: no reinsertion or identity-swap moves");
      if (d.charged) return fail(GRASPA_ERR_UNSUPPORTED, "flexible chains must be uncharged");
    }
    if (sp.p_swap > 1e-10) h.valid[h.n_valid++] = c;
  }
  if (h.n_valid == 0) return fail(GRASPA_ERR_ARG, "no species with a move probability");
  h.K = p->n_trial_pos; h.O = p->n_trial_orient; h.bias = p->vdw_real_bias;
  h.use_max_step = p->use_max_step; h.max_step = p->max_step_per_cycle;
  h.beta = p->beta; h.pressure = p->pressure;
  h.seed = in.seed; h.stream = in.stream;
  h.has_tail = p->has_tail;
  h.trace_cap = bt->trace_cap;
  h.cyc_cap = bt->cyc_cap;
  h.n_block = in.n_block > 0 ? in.n_block : 5;
  h.trace = bt->d_trace;
  h.cyc = bt->d_cyc ? bt->d_cyc + (size_t)i * bt->cyc_cap : nullptr;
  h.blk = bt->d_blk + i;
  h.tail = bt->d_tail + (size_t)i * kMaxTypes * kMaxTypes;
  h.use_tail = bt->d_use_tail + (size_t)i * kMaxTypes * kMaxTypes;
  if (p->has_tail) {
    if (hip_check(hipMemcpy(bt->d_tail + (size_t)i * kMaxTypes * kMaxTypes, p->tail, (size_t)T * T * sizeof(double),
                            hipMemcpyHostToDevice), "tail upload"))
      return GRASPA_ERR_HIP;
    if (hip_check(hipMemcpy(bt->d_use_tail + (size_t)i * kMaxTypes * kMaxTypes, p->use_tail,
                            (size_t)T * T * sizeof(int), hipMemcpyHostToDevice), "tail upload"))
      return GRASPA_ERR_HIP;
  }
  BState st;
  std::memset(&st, 0, sizeof(st));
  st.total_mol = 1;
  for (int c = 1; c < ncomp; ++c) {
    st.n_mol[c] = b->comps[c].n_mol;
    st.total_mol += b->comps[c].n_mol;
    for (int d = 0; d < 3; ++d) {
      st.max_trans[c][d] = p->species[c - 1].max_translation[d];
      st.max_rot[c][d] = p->species[c - 1].max_rotation[d];
    }
  }
  for (int t = 0; t < GRASPA_MAX_PSEUDO; ++t) st.n_pseudo[t] = p->n_pseudo[t];
  st.blk_size = 1;
  if (hip_check(hipMemset(bt->d_blk + i, 0, sizeof(graspa_block_stats)), "block statistics")) return GRASPA_ERR_HIP;
  hipSetDevice(b->device);
  if (hip_check(hipStreamSynchronize(b->stream), "box stream")) return GRASPA_ERR_HIP;
  // both S(k) buffers start equal, so entries a move never writes (inactive
  // k-vectors) agree whichever buffer is current
  if (b->bc.use_charges && b->nvec > 0 &&
      hip_check(hipMemcpy(b->d_sk_stage, b->d_sk_ads, b->nvec * sizeof(double2), hipMemcpyDeviceToDevice), "sk copy"))
    return GRASPA_ERR_HIP;
  bt->hb[i] = h;
  bt->hs[i] = st;
  bt->boxes[i] = b;
  bt->us_per_move[i] = 0.0;
  bt->seen_moves[i] = 0;
  bt->seen_busy[i] = 0;
  bt->lds = std::max(bt->lds, need);
  bt->cubic = 1;
  for (int k = 0; k < bt->n; ++k)
    if (bt->boxes[k]) bt->cubic &= bt->hb[k].bc.cubic ? 1 : 0;
  return GRASPA_OK;
}

template <bool MIXED, bool CUBIC>
int launch_batch(graspa_batch* bt, int max_moves)
{
  auto kern = batch_mc_kernel<MIXED, CUBIC>;
  // all the LDS a workgroup can have (one box workgroup fills a CU anyway):
  // what the tables leave holds the framework when it fits
  hipFuncAttributes fa;
  size_t stat = 32768;
  if (hipFuncGetAttributes(&fa, (const void*)kern) == hipSuccess) stat = fa.sharedSizeBytes;
  const size_t avail = bt->max_lds > (int)stat ? (((size_t)bt->max_lds - stat) & ~(size_t)15) : 0;
  if (avail < bt->lds) return fail(GRASPA_ERR_UNSUPPORTED, "batch workspace exceeds LDS");
  const size_t lds = avail;
  (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (!bt->ev0 && hip_check(hipEventCreate(&bt->ev0), "batch event")) return GRASPA_ERR_HIP;
  if (!bt->ev1 && hip_check(hipEventCreate(&bt->ev1), "batch event")) return GRASPA_ERR_HIP;
  hipEventRecord(bt->ev0, bt->stream);
  const int64_t* bud = nullptr;
  if (bt->slice_us > 0.0) {
    // budgets for a launch of about slice_us: each box's moves at its
    // measured time per move (a new box starts from 20 us per move)
    for (int i = 0; i < bt->n; ++i) {
      const double c = bt->us_per_move[i] > 0.0 ? bt->us_per_move[i] : 20.0;
      bt->hbud[i] = std::max<int64_t>(1, std::min<int64_t>(max_moves, (int64_t)(bt->slice_us / c)));
    }
    if (hip_check(hipMemcpyAsync(bt->d_bud, bt->hbud.data(), bt->n * sizeof(int64_t), hipMemcpyHostToDevice,
                                 bt->stream), "batch budgets"))
      return GRASPA_ERR_HIP;
    bud = bt->d_bud;
  }
  hipLaunchKernelGGL(kern, dim3(bt->n), dim3(kBT), lds, bt->stream, bt->d_box, bt->d_state, max_moves, bud, (int)lds);
  if (hip_check(hipGetLastError(), "batch_mc_kernel launch")) return GRASPA_ERR_HIP;
  hipEventRecord(bt->ev1, bt->stream);
  bt->in_flight = true;
  return GRASPA_OK;
}

// Waits for the launch in flight and adds its kernel time.
int finish_launch(graspa_batch* bt)
{
  if (!bt->in_flight) return GRASPA_OK;
  bt->in_flight = false;
  if (hip_check(hipEventSynchronize(bt->ev1), "batch_mc_kernel")) return GRASPA_ERR_HIP;
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, bt->ev0, bt->ev1) == hipSuccess) bt->kernel_ms += ms;
  bt->launches++;
  return GRASPA_OK;
}

int launch_once(graspa_batch* bt, int max_moves)
{
  if (bt->in_flight) return fail(GRASPA_ERR_STATE, "a launch of the batch is in flight (graspa_batch_wait)");
  if (bt->precision == GRASPA_MIXED)
    return bt->cubic ? launch_batch<true, true>(bt, max_moves) : launch_batch<true, false>(bt, max_moves);
  return bt->cubic ? launch_batch<false, true>(bt, max_moves) : launch_batch<false, false>(bt, max_moves);
}

}  // namespace

extern "C" {

int graspa_batch_create(const graspa_batch_box* boxes, int n_boxes, int trace_cap, int cycle_cap, graspa_batch** out)
{
  if (n_boxes < 1 || !out || trace_cap < 0 || cycle_cap < 0) return fail(GRASPA_ERR_ARG, "bad arguments");
  *out = nullptr;
  graspa_batch* bt = new graspa_batch();
  bt->n = n_boxes;
  bt->trace_cap = trace_cap;
  bt->cyc_cap = cycle_cap;
  bt->boxes.assign(n_boxes, nullptr);
  bt->skc.assign(n_boxes, nullptr);
  bt->hb.assign(n_boxes, BBox{});
  bt->hs.assign(n_boxes, BState{});
  bt->hbud.assign(n_boxes, 0);
  bt->us_per_move.assign(n_boxes, 0.0);
  bt->seen_moves.assign(n_boxes, 0);
  bt->seen_busy.assign(n_boxes, 0);
  for (int i = 0; i < n_boxes; ++i) {
    bt->hs[i].error = 0;
    bt->hs[i].target = 0;   // empty slots never run
  }
  // device and precision of the batch: those of the first box given
  bool have = false;
  for (int i = 0; boxes && i < n_boxes && !have; ++i)
    if (boxes[i].box) {
      bt->device = boxes[i].box->device;
      bt->precision = boxes[i].box->precision;
      have = true;
    }
  hipSetDevice(bt->device);
  int rc = hip_check(hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking), "batch stream");
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, bt->device) == hipSuccess) {
      bt->max_lds = (int)prop.sharedMemPerBlock;
      int v = 0;
      if (hipDeviceGetAttribute(&v, hipDeviceAttributeSharedMemPerBlockOptin, bt->device) == hipSuccess &&
          v > bt->max_lds)
        bt->max_lds = v;
    }
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, bt->device) == hipSuccess && khz > 0)
      bt->tick_hz = 1000.0 * khz;
  }
  const size_t nt = (size_t)n_boxes * kMaxTypes * kMaxTypes;
  rc = rc ? rc : hip_check(hipMalloc(&bt->d_tail, nt * sizeof(double)), "hipMalloc batch");
  rc = rc ? rc : hip_check(hipMalloc(&bt->d_use_tail, nt * sizeof(int)), "hipMalloc batch");
  if (!rc && trace_cap > 0)
    rc = hip_check(hipMalloc(&bt->d_trace, (size_t)n_boxes * trace_cap * sizeof(graspa_trace)), "hipMalloc batch");
  if (!rc && cycle_cap > 0)
    rc = hip_check(hipMalloc(&bt->d_cyc, (size_t)n_boxes * cycle_cap * sizeof(float)), "hipMalloc batch");
  rc = rc ? rc : hip_check(hipMalloc(&bt->d_blk, n_boxes * sizeof(graspa_block_stats)), "hipMalloc batch");
  rc = rc ? rc : hip_check(hipMemset(bt->d_blk, 0, n_boxes * sizeof(graspa_block_stats)), "block statistics");
  rc = rc ? rc : hip_check(hipMalloc(&bt->d_box, n_boxes * sizeof(BBox)), "hipMalloc batch");
  rc = rc ? rc : hip_check(hipMalloc(&bt->d_state, n_boxes * sizeof(BState)), "hipMalloc batch");
  rc = rc ? rc : hip_check(hipMalloc(&bt->d_bud, n_boxes * sizeof(int64_t)), "hipMalloc batch");
  for (int i = 0; boxes && i < n_boxes && !rc; ++i)
    if (boxes[i].box) rc = fill_slot(bt, i, boxes[i]);
  rc = rc ? rc : upload_batch(bt);
  if (rc) { free_batch(bt); return rc; }
  *out = bt;
  return GRASPA_OK;
}

int graspa_batch_set_box(graspa_batch* bt, int slot, const graspa_batch_box* box)
{
  if (!bt || slot < 0 || slot >= bt->n) return fail(GRASPA_ERR_ARG, "bad slot");
  if (int rc = finish_launch(bt)) return rc;
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  if (!box || !box->box) {
    bt->boxes[slot] = nullptr;
    bt->hb[slot] = BBox{};
    bt->hs[slot] = BState{};
    hipFree(bt->skc[slot]);
    bt->skc[slot] = nullptr;
  } else {
    const bool first = std::none_of(bt->boxes.begin(), bt->boxes.end(), [](graspa_box* x) { return x != nullptr; });
    if (first) {
      bt->device = box->box->device;
      bt->precision = box->box->precision;
    }
    if (int rc = fill_slot(bt, slot, *box)) return rc;
  }
  return upload_batch(bt);
}

// A production run starts: blocks of n_cycles / n_block cycles
// (axpy.cu:500-502), statistics cleared.
static int start_production(graspa_batch* bt, int slot, int n_cycles)
{
  BState& s = bt->hs[slot];
  s.prod_first = s.cycle;
  s.blk_size = std::max(1, n_cycles / bt->hb[slot].n_block);
  return hip_check(hipMemset(bt->d_blk + slot, 0, sizeof(graspa_block_stats)), "block statistics");
}

int graspa_batch_set_cycles(graspa_batch* bt, int slot, int n_cycles, int production)
{
  if (!bt || slot < 0 || slot >= bt->n || n_cycles < 0) return fail(GRASPA_ERR_ARG, "bad arguments");
  if (int rc = finish_launch(bt)) return rc;
  if (!bt->boxes[slot]) return fail(GRASPA_ERR_STATE, "empty slot");
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  BState& s = bt->hs[slot];
  s.target = s.cycle + n_cycles;
  s.production = production ? 1 : 0;
  s.trace_n = 0;
  s.cyc_first = s.cycle;
  if (production)
    if (int rc = start_production(bt, slot, n_cycles)) return rc;
  return hip_check(hipMemcpy(&bt->d_state[slot], &s, sizeof(BState), hipMemcpyHostToDevice), "batch state upload");
}

int graspa_batch_set_time_slice(graspa_batch* bt, double us)
{
  if (!bt || !(us >= 0.0)) return fail(GRASPA_ERR_ARG, "bad arguments");
  bt->slice_us = us;
  return GRASPA_OK;
}

int graspa_batch_launch_async(graspa_batch* bt, int moves_per_launch)
{
  if (!bt) return fail(GRASPA_ERR_ARG, "null batch");
  hipSetDevice(bt->device);
  return launch_once(bt, moves_per_launch > 0 ? moves_per_launch : 4096);
}

int graspa_batch_launch(graspa_batch* bt, int moves_per_launch, int* n_running)
{
  if (int rc = graspa_batch_launch_async(bt, moves_per_launch)) return rc;
  return graspa_batch_wait(bt, n_running);
}

int graspa_batch_wait(graspa_batch* bt, int* n_running)
{
  if (!bt) return fail(GRASPA_ERR_ARG, "null batch");
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  int run = 0;
  for (int i = 0; i < bt->n; ++i) {
    const BState& s = bt->hs[i];
    if (!bt->boxes[i]) continue;
    if (s.error) return fail(s.error, "a box of the batch stopped on an error (molecule capacity)");
    if (s.cycle < s.target) ++run;
  }
  if (n_running) *n_running = run;
  return GRASPA_OK;
}

void graspa_batch_destroy(graspa_batch* bt) { free_batch(bt); }

int graspa_batch_run(graspa_batch* bt, int n_cycles, int production, int moves_per_launch, int64_t* moves)
{
  if (!bt || n_cycles < 0) return fail(GRASPA_ERR_ARG, "bad arguments");
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  int64_t before = 0;
  for (int i = 0; i < bt->n; ++i) {
    BState& s = bt->hs[i];
    if (!bt->boxes[i]) continue;
    if (s.error) return fail(s.error, "a box of the batch stopped on an error (molecule capacity)");
    s.target = s.cycle + n_cycles;
    s.production = production ? 1 : 0;
    s.trace_n = 0;
    s.cyc_first = s.cycle;
    if (production)
      if (int rc = start_production(bt, i, n_cycles)) return rc;
    before += s.move;
  }
  if (int rc = hip_check(hipMemcpyAsync(bt->d_state, bt->hs.data(), bt->n * sizeof(BState), hipMemcpyHostToDevice,
                                        bt->stream), "batch state upload"))
    return rc;
  const int mpl = moves_per_launch > 0 ? moves_per_launch : 4096;
  int64_t last = before;
  for (;;) {
    int rc = launch_once(bt, mpl);
    if (rc) return rc;
    if ((rc = download_states(bt))) return rc;
    bool done = true;
    int err = 0;
    for (int i = 0; i < bt->n; ++i) {
      const BState& s = bt->hs[i];
      if (!bt->boxes[i]) continue;
      if (s.error) err = s.error;
      else if (s.cycle < s.target) done = false;
    }
    if (err) return fail(err, "a box of the batch stopped on an error (molecule capacity)");
    if (done) break;
    int64_t now = 0;
    for (int i = 0; i < bt->n; ++i) now += bt->hs[i].move;
    if (now == last) return fail(GRASPA_ERR_STATE, "batch launch made no progress");
    last = now;
  }
  int64_t after = 0;
  for (int i = 0; i < bt->n; ++i) after += bt->hs[i].move;
  if (moves) *moves = after - before;
  return GRASPA_OK;
}

int graspa_batch_block_stats(graspa_batch* bt, int i, graspa_block_stats* out)
{
  if (!bt || i < 0 || i >= bt->n || !out) return fail(GRASPA_ERR_ARG, "bad arguments");
  if (!bt->boxes[i]) return fail(GRASPA_ERR_STATE, "empty slot");
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  if (hip_check(hipMemcpy(out, bt->d_blk + i, sizeof(graspa_block_stats), hipMemcpyDeviceToHost), "block statistics"))
    return GRASPA_ERR_HIP;
  out->block_size = bt->hs[i].blk_size;
  out->n_block = bt->hb[i].n_block;
  return GRASPA_OK;
}

int graspa_batch_box_stats(graspa_batch* bt, int i, graspa_energy* running, int* n_mol, int64_t* counters,
                           double* widom_sum_w, double* widom_n, double* load_sum, double* load_sum2,
                           int64_t* n_load_cycles, int64_t* moves, int64_t* cycles)
{
  if (!bt || i < 0 || i >= bt->n || !bt->boxes[i]) return fail(GRASPA_ERR_ARG, "bad box index or empty slot");
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  const BState& s = bt->hs[i];
  const int ncomp = bt->hb[i].n_comp;
  if (running) std::memcpy(running, s.running, sizeof(graspa_energy));
  for (int c = 1; c < ncomp; ++c) {
    if (n_mol) n_mol[c - 1] = s.n_mol[c];
    if (counters)
      for (int k = 0; k < GRASPA_N_COUNTERS; ++k) counters[GRASPA_N_COUNTERS * (c - 1) + k] = s.cnt[c][k];
    if (widom_sum_w) widom_sum_w[c - 1] = s.widom_w[c];
    if (widom_n) widom_n[c - 1] = s.widom_n[c];
    if (load_sum) load_sum[c - 1] = s.load_sum[c];
    if (load_sum2) load_sum2[c - 1] = s.load_sum2[c];
  }
  if (n_load_cycles) *n_load_cycles = s.load_n;
  if (moves) *moves = s.move;
  if (cycles) *cycles = s.cycle;
  return GRASPA_OK;
}

int graspa_batch_max_steps(graspa_batch* bt, int i, double* max_translation, double* max_rotation)
{
  if (!bt || i < 0 || i >= bt->n || !bt->boxes[i]) return fail(GRASPA_ERR_ARG, "bad box index or empty slot");
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  const BState& s = bt->hs[i];
  for (int c = 1; c < bt->hb[i].n_comp; ++c)
    for (int d = 0; d < 3; ++d) {
      if (max_translation) max_translation[3 * (c - 1) + d] = s.max_trans[c][d];
      if (max_rotation) max_rotation[3 * (c - 1) + d] = s.max_rot[c][d];
    }
  return GRASPA_OK;
}

int graspa_batch_cycle_times(graspa_batch* bt, int i, double* seconds, int max_n)
{
  if (!bt || i < 0 || i >= bt->n || max_n < 0 || (max_n > 0 && !seconds)) return fail(GRASPA_ERR_ARG, "bad arguments");
  if (!bt->d_cyc) return 0;
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  const BState& s = bt->hs[i];
  const int64_t first = std::max(s.cyc_first, s.cycle - (int64_t)bt->cyc_cap);
  const int n = (int)std::min<int64_t>(max_n, s.cycle - first);
  std::vector<float> ring(bt->cyc_cap);
  if (hip_check(hipMemcpy(ring.data(), bt->d_cyc + (size_t)i * bt->cyc_cap, bt->cyc_cap * sizeof(float),
                          hipMemcpyDeviceToHost), "cycle times copy"))
    return GRASPA_ERR_HIP;
  for (int k = 0; k < n; ++k) seconds[k] = (double)ring[(first + k) % bt->cyc_cap] / bt->tick_hz;
  return n;
}

int graspa_batch_trace(graspa_batch* bt, int i, graspa_trace* out, int max_n)
{
  if (!bt || i < 0 || i >= bt->n || max_n < 0 || (max_n > 0 && !out)) return fail(GRASPA_ERR_ARG, "bad arguments");
  if (!bt->d_trace) return 0;
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  const int n = (int)std::min<int64_t>(std::min<int64_t>(max_n, bt->trace_cap), bt->hs[i].trace_n);
  if (n > 0 && hip_check(hipMemcpy(out, bt->d_trace + (size_t)i * bt->trace_cap, n * sizeof(graspa_trace),
                                   hipMemcpyDeviceToHost), "trace copy"))
    return GRASPA_ERR_HIP;
  return n;
}

int graspa_batch_progress(graspa_batch* bt, int64_t* moves, int64_t* cycles, int64_t* targets)
{
  if (!bt) return fail(GRASPA_ERR_ARG, "null batch");
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  for (int i = 0; i < bt->n; ++i) {
    const bool on = bt->boxes[i] != nullptr;
    if (moves) moves[i] = on ? bt->hs[i].move : 0;
    if (cycles) cycles[i] = on ? bt->hs[i].cycle : 0;
    if (targets) targets[i] = on ? bt->hs[i].target : 0;
  }
  return GRASPA_OK;
}

int graspa_batch_busy(graspa_batch* bt, double* seconds)
{
  if (!bt || !seconds) return fail(GRASPA_ERR_ARG, "bad arguments");
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  for (int i = 0; i < bt->n; ++i) seconds[i] = bt->boxes[i] ? (double)bt->hs[i].busy / bt->tick_hz : 0.0;
  return GRASPA_OK;
}

int graspa_batch_kernel_time(graspa_batch* bt, double* total_ms, int64_t* launches)
{
  if (!bt) return fail(GRASPA_ERR_ARG, "null batch");
  if (total_ms) *total_ms = bt->kernel_ms;
  if (launches) *launches = bt->launches;
  return GRASPA_OK;
}

int graspa_batch_sync(graspa_batch* bt)
{
  if (!bt) return fail(GRASPA_ERR_ARG, "null batch");
  hipSetDevice(bt->device);
  if (int rc = download_states(bt)) return rc;
  for (int i = 0; i < bt->n; ++i) {
    graspa_box* b = bt->boxes[i];
    if (!b) continue;
    BState& s = bt->hs[i];
    const int cap = b->guest_cap;
    std::vector<double> g(4 * (size_t)std::max(cap, 1));
    std::vector<int> gt(std::max(cap, 1)), gm(std::max(cap, 1));
    if (cap > 0) {
      if (hip_check(hipMemcpy(g.data(), b->d_g, 4 * (size_t)cap * sizeof(double), hipMemcpyDeviceToHost), "sync"))
        return GRASPA_ERR_HIP;
      if (hip_check(hipMemcpy(gt.data(), b->d_g_type, (size_t)cap * sizeof(int), hipMemcpyDeviceToHost), "sync"))
        return GRASPA_ERR_HIP;
      if (hip_check(hipMemcpy(gm.data(), b->d_g_molid, (size_t)cap * sizeof(int), hipMemcpyDeviceToHost), "sync"))
        return GRASPA_ERR_HIP;
    }
    for (size_t c = 1; c < b->comps.size(); ++c) {
      Component& cc = b->comps[c];
      cc.n_mol = s.n_mol[c];
      const int na = std::max(cc.n_mol, 1) * cc.molsize;
      for (int a = 0; a < na; ++a) {
        const int j = cc.base + a;
        cc.pos[3 * a] = g[j];
        cc.pos[3 * a + 1] = g[cap + j];
        cc.pos[3 * a + 2] = g[2 * (size_t)cap + j];
        cc.q[a] = g[3 * (size_t)cap + j];
        cc.type[a] = gt[j];
        cc.molid[a] = gm[j];
      }
      // refreshes the fixed-point copy the screening kernels read
      if (int rc = upload_guest_atoms(b, (int)c, 0, na)) return rc;
    }
    const BBox& h = bt->hb[i];
    if (h.n_active > 0) {
      // the current adsorbate S(k) of the active k-vectors back into the box's arrays
      const int na = h.n_active, nv = b->nvec;
      std::vector<int4> ka(na);
      std::vector<double2> ads(nv), c(na);
      if (hip_check(hipMemcpy(ka.data(), b->d_kact, na * sizeof(int4), hipMemcpyDeviceToHost), "sync") ||
          hip_check(hipMemcpy(ads.data(), b->d_sk_ads, nv * sizeof(double2), hipMemcpyDeviceToHost), "sync") ||
          hip_check(hipMemcpy(c.data(), h.skc[s.sk_cur], na * sizeof(double2), hipMemcpyDeviceToHost), "sync"))
        return GRASPA_ERR_HIP;
      for (int j = 0; j < na; ++j) ads[ka[j].w] = c[j];
      if (hip_check(hipMemcpy(b->d_sk_ads, ads.data(), nv * sizeof(double2), hipMemcpyHostToDevice), "sync") ||
          hip_check(hipMemcpy(b->d_sk_stage, ads.data(), nv * sizeof(double2), hipMemcpyHostToDevice), "sync"))
        return GRASPA_ERR_HIP;
    }
    b->sk_staged = false;
    if (hip_check(hipStreamSynchronize(b->stream), "sync")) return GRASPA_ERR_HIP;
  }
  return GRASPA_OK;
}

}  // extern "C"
