// widom_kernel.hip -- batched Widom / CBMC test insertions (the throughput path
// of the GCMC energy engine on gfx950).
//
// A batch runs as a short pipeline of small kernels on the box's stream,
// linked by compact work lists in HBM, so that each kernel keeps few registers
// live and runs at high occupancy:
//   A  (one lane per first-bead trial)  the trial's position from the
//      insertion's Philox stream and its overlap screen; an insertion with no
//      trial left is finished here (W = 0, most insertions in a dense
//      framework), the others go to a list of live insertions and their
//      surviving trial sites to list 1;
//   B  (one wave per trial site) full pair energies (LJ + real-space Coulomb,
//      FP64 accumulation) of every site in the list, with the force-field rows,
//      the real-space table and the framework staged in LDS; exact overlap
//      flags;
//   C  (one wave per live insertion)  first-bead selection; O orientations
//      of the template about the selected bead, one lane per orientation
//      site, and their screen; survivors to list 2;
//   D  = B on list 2;
//   E  (one wave per live insertion)  orientation selection, Fourier-space energy
//      of the selected molecule against the resident structure factors, W and
//      the per-insertion record.
// Two small kernels then sum the records in index order (deterministic sums).
// Every random number of insertion i comes from the counter-based stream
// (seed, stream, first_index + i), so a result depends on nothing but i.
//
// The overlap screen: FP32 squared distances from fixed-point coordinates
// against per-(site type, atom type) thresholds r_ov^2 below which the pair
// energy certainly exceeds OverlapCriteria (or r^2 < 0.01), over the atoms of
// the 27 grid cells around the trial site (cells at least r_ov wide, so this
// finds every overlapping atom); a trial with a hit is one the reference flags
// (VDW_Coulomb.cu:1299-1300) and leaves out of the selection
// (Host_sum_Widom_HGGG_SEPARATE, mc_widom.h:41-87), so its energies are not
// computed.
// The semantics are Insertion_Body (mc_swap_utilities.h:3-133) with
// Widom_Move_FirstBead_PARTIAL / Widom_Move_Chain_PARTIAL (mc_widom.h:305-614).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "device_math.hpp"
#include "engine.hpp"
#include "move_math.hpp"

namespace graspa {

constexpr int kMaxTrials = 32;   // K and O (trial masks are 32-bit)
constexpr int kMaxMol = 8;       // sites per molecule
constexpr int kMaxPts = 64;      // trial sites of one screen: K, or O x (M-1)
constexpr int kSums = 6;         // W, W^2, W dE, n, n first bead placed, n complete
constexpr int kRedBlock = 256;
constexpr int kRedPer = 8;       // records per thread in the first reduction pass
constexpr int kInsWaves = 16;    // waves per workgroup of the chain kernel (one live insertion each)
constexpr int kFinWaves = 4;     // waves per workgroup of the finish kernel
constexpr int kFinMaxPer = 8;    // finish workgroups per placed list at most
constexpr int kFbWaves = 16;     // waves per workgroup of the first-bead kernel

// Work lists are appended per workgroup of the screening kernels: workgroup g
// of a grid of G owns insertions [g*chunk, (g+1)*chunk) and the sub-list at
// g*chunk*cap_per_insertion; its length goes to counts[g].
constexpr int kMaxScreenGroups = 1024;
constexpr int kChainScreenMinAtoms = 4096;   // frameworks from this size screen orientation sites
constexpr int kMaxTunePerCu = 8;       // workgroups per CU the autotuner tries at most
constexpr int kMaxEvalGroups = 4096;   // workgroups of one pair-energy launch (work-count slots)

// State of one insertion between the kernels.
struct InsState {
  double fb[3];                  // selected first-bead position
  double wfb, ev, er;            // first-bead Rosenbluth weight, its LJ and real-space energies
  int sel;                       // selected first-bead trial, -1 if none
  int ok;                        // first bead placed
};

// One trial site to evaluate.
struct TrialPt {
  double x, y, z;
  int site;                      // molecule site (0 = first bead, 1 + a = chain atom a)
  int slot;                      // index of its energy slot
  int ins;                       // insertion (trial-mask word)
  int bit;                       // trial bit in that word
};

struct WidomArgs {
  BoxConst bc;
  const double* fw;        // SoA x,y,z,q [4][n_fw], Morton order of the screen cells
  const int* fw_type;      // same order
  const uint4* fw_fixT;    // fixed-point x,y,z + type (scan without a cell grid)
  int n_fw;
  int use_cells;
  int screen_chain;             // cell screen of the orientation sites (large frameworks only)
  int iso;                      // cubic cell with equal edges: the fast form's pair kernel works in cell fractions
  int lj_general;               // the fast pair loop's general form: some LJ shift is nonzero, or a
                                // molecule site's LJ with some framework type at r = 0.1 A stays
                                // below twice OverlapCriteria (the r^2 < 0.01 test is then needed)
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
  const int64_t* first_dev;  // when set, the first index is read here (graph replays)
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
  // pipeline buffers
  uint32_t* mask1;          // [n_insert] first-bead trials with an overlap
  double2* e1;              // [n_insert * K] first-bead (LJ, real) energies
  TrialPt* list1;
  InsState* state;
  uint32_t* mask2;          // [n_insert] orientations with an overlap
  double2* e2;              // [n_insert * O * chain] per chain site
  TrialPt* list2;
  unsigned* counts1;        // [groups] sub-list lengths of list 1
  unsigned* counts2;        // [groups] sub-list lengths of list 2
  unsigned* live_counts;    // [groups] lengths of the live sub-lists
  const uint32_t* dmap[kMaxMol];   // certain-overlap voxel map of each molecule site's type (or null)
  const uint32_t* fmap0;           // certainly-free voxel map of the first bead's type (or null)
  int dnv[3];               // its voxels per cell axis
  int* live;                // insertions with a first-bead trial left after the screen, sub-list
                            // of workgroup g at g * chunk
  int* placed;              // insertions whose first bead was placed, sub-list of workgroup g at
                            // g * chunk (in list 1's memory, free once list 1 is evaluated)
  unsigned* placed_counts;  // [groups] their counts (in counts1's memory, likewise)
  unsigned long long* stat_part;   // per-workgroup work counts (screened, evaluated pairs)
  int groups, chunk;        // screening grid: workgroups and insertions per workgroup
  // outputs
  graspa_widom_result* res;   // may be null
  double* rec_w;              // [n_insert] W
  double* rec_e;              // [n_insert] dE
  int* rec_f;                 // [n_insert] bit 0 first bead placed, bit 1 complete
  unsigned long long* stat;   // screened pairs, evaluated pairs (sums of stat_part)
  // LDS layout (bytes from the start of the dynamic area)
  int lds_thr, lds_tab;                 // screening kernels
  int lds_ffl, lds_coul, lds_fw;        // pair-energy kernel (lds_coul is 0)
};

// First insertion index of the batch.
__device__ inline int64_t first_of(const WidomArgs& A) { return A.first_dev ? *A.first_dev : A.first_index; }

__device__ inline double draw(const WidomArgs& A, uint64_t idx, uint32_t d)
{
  return philox_uniform(A.seed, A.stream, idx, d);
}

// Uniforms d0, d0 + 1, d0 + 2 of insertion idx (philox_uniform's layout: draw d
// is half d & 1 of Philox block d >> 1; the three draws span two blocks).
__device__ inline void draw3(const WidomArgs& A, uint64_t idx, int d0, double u[3])
{
  uint32_t c0[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)(d0 >> 1), A.stream};
  uint32_t c1[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)((d0 >> 1) + 1), A.stream};
  philox4x32_10(c0, (uint32_t)A.seed, (uint32_t)(A.seed >> 32));
  philox4x32_10(c1, (uint32_t)A.seed, (uint32_t)(A.seed >> 32));
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int d = d0 + k;
    const uint32_t* c = ((d >> 1) == (d0 >> 1)) ? c0 : c1;
    const uint32_t a = (d & 1) ? c[2] : c[0], b = (d & 1) ? c[3] : c[1];
    u[k] = ((double)((((uint64_t)a << 32) | b) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
}

// First-bead trial position t: diag(cell) * U^3 (get_random_trial_position,
// mc_widom.h:122-213), draws 3t .. 3t+2.
__device__ inline void first_bead_pos(const WidomArgs& A, uint64_t idx, int t, double p[3])
{
  p[0] = A.bc.cell[0] * draw(A, idx, 3 * t);
  p[1] = A.bc.cell[4] * draw(A, idx, 3 * t + 1);
  p[2] = A.bc.cell[8] * draw(A, idx, 3 * t + 2);
}

// Chain atom a of orientation o about the first bead fb (get_random_trial_orientation,
// mc_widom.h:215-303): the template vector rotated by the quaternion from draws
// 3K+1+3o .. 3K+3+3o.
__device__ inline void orient_pos(const WidomArgs& A, uint64_t idx, int o, int a, const double fb[3], double p[3])
{
  const double u = draw(A, idx, 3 * A.K + 1 + 3 * o);
  const double v = draw(A, idx, 3 * A.K + 2 + 3 * o);
  const double w = draw(A, idx, 3 * A.K + 3 + 3 * o);
  double vec[3] = {A.tvec[a][0], A.tvec[a][1], A.tvec[a][2]};
  quat_rotate(vec, u, v, w);
  p[0] = fb[0] + vec[0];
  p[1] = fb[1] + vec[1];
  p[2] = fb[2] + vec[2];
}

// Overlap screen of one trial site per lane: FP32 squared distances from the
// fixed-point site P to the framework atoms of the 27 grid cells around it
// (the cells are at least as wide as the largest overlap distance, so they
// hold every framework atom that can overlap the site) against the site's
// thresholds th[atom type]; stops at the first hit.  Guests are not screened:
// their overlaps are found by the exact evaluation.
template <bool CUBIC>
__device__ inline bool screen_row(const WidomArgs& A, const uint4* __restrict__ tab, const float* __restrict__ th,
                                  const uint4 P, const int r, unsigned long long& st)
{
  // row r (0..8) of the 3 x 3 x 3 cells around P's cell: dz = r / 3 - 1, dy = r % 3 - 1
  const int C = A.cell_cap, n0 = A.cell_n[0], n1 = A.cell_n[1], n2 = A.cell_n[2];
  const int cx = (int)(((uint64_t)P.x * (uint64_t)n0) >> 32);
  const int cy = (int)(((uint64_t)P.y * (uint64_t)n1) >> 32);
  const int cz = (int)(((uint64_t)P.z * (uint64_t)n2) >> 32);
  const int xm = cx == 0 ? n0 - 1 : cx - 1, xp = cx == n0 - 1 ? 0 : cx + 1;
  int z = cz + r / 3 - 1;
  z = z < 0 ? z + n2 : (z >= n2 ? z - n2 : z);
  int y = cy + r % 3 - 1;
  y = y < 0 ? y + n1 : (y >= n1 ? y - n1 : y);
  const int row = __mul24(n0, y + __mul24(n1, z));
  const int cb[3] = {__mul24(row + xm, C), __mul24(row + cx, C), __mul24(row + xp, C)};
  // a row of three cells per step, up to four slots each, loaded together
  // (slots fill from the front, empty ones have type ~0u)
  for (int k0 = 0; k0 < C; k0 += 4) {
    uint4 at[12];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) at[4 * c + j] = (k0 + j < C) ? tab[cb[c] + k0 + j] : make_uint4(0u, 0u, 0u, ~0u);
    bool hit = false;
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      if (at[q].w != ~0u) {
        ++st;
        hit = hit || (r2_fixed32<CUBIC>(at[q].x, at[q].y, at[q].z, P.x, P.y, P.z, A.bc.h32) < th[at[q].w]);
      }
    }
    if (hit) return true;
  }
  return false;
}

// The whole 3 x 3 x 3 neighbourhood of P on one lane, row by row.
template <bool CUBIC>
__device__ inline bool screen_lane(const WidomArgs& A, const uint4* __restrict__ tab, const float* __restrict__ th,
                                   const uint4 P, unsigned long long& st)
{
  for (int r = 0; r < 9; ++r)
    if (screen_row<CUBIC>(A, tab, th, P, r, st)) return true;
  return false;
}

// Bit of a voxel map (certain overlap, or certainly free) at the voxel of the
// fixed-point site P.
__device__ inline bool voxel_bit(const uint32_t* __restrict__ dm, const int* nv, const uint4 P)
{
  const uint32_t x = (uint32_t)(((uint64_t)P.x * (uint64_t)nv[0]) >> 32);
  const uint32_t y = (uint32_t)(((uint64_t)P.y * (uint64_t)nv[1]) >> 32);
  const uint32_t z = (uint32_t)(((uint64_t)P.z * (uint64_t)nv[2]) >> 32);
  const uint32_t lin = x + (uint32_t)nv[0] * (y + (uint32_t)nv[1] * z);
  return (dm[lin >> 5] >> (lin & 31u)) & 1u;
}

// SelectTrialPosition / CBMC_FirstBead_Finish / Widom_Move_Chain_PARTIAL
// (mc_widom.h:14-39, 305-383, 574-614): lane t < n holds trial t's LJ (ev) and
// real-space (er) sums; ovm marks overlapping trials.  Survivors, log-sum-exp
// shift, cumulative pick with the uniform u in trial order, Rosenbluth sum
// averaged over n, and the LJ-only bias correction.  Returns the selected
// trial (or -1) and W, both uniform.
__device__ inline void wave_select(double ev, double er, uint32_t ovm, int n, double u, double beta, int bias,
                                   int& sel_out, double& W_out)
{
  const int lane = threadIdx.x & 63;
  const bool alive = lane < n && !((ovm >> lane) & 1u);
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

// wave_select for the insertion of this lane's 16-lane row (its trials in
// lanes 0 .. n-1 of the row, n <= 16): four insertions per wave.  The same
// arithmetic in the same order: the row maximum by the same maximum steps, the
// ordered sums over the survivors left to right by a scan along the row (after
// step t, lane t holds ((s_0 + s_1) + ...) + s_t; screened trials add 0.0),
// and the cumulative pick is the first survivor whose running sum is not below
// u times the total, else the last survivor, as the sequential loop stops.
__device__ inline void row_select(double ev, double er, bool alive, int n, double u, double beta, int bias,
                                  int& sel_out, double& W_out)
{
  const int lane = threadIdx.x & 63, rl = lane & 15, rb0 = lane & 48;
  double lw = -INFINITY;
  if (alive) {
    double tot = ev;
    if (bias) tot += er;
    lw = -beta * tot;
  }
  double mx = lw;
  mx = fmax(mx, dpp_d<0xB1>(mx));         // quad_perm [1,0,3,2]
  mx = fmax(mx, dpp_d<0x4E>(mx));         // quad_perm [2,3,0,1]
  mx = fmax(mx, dpp_d<0x124>(mx));        // row_ror:4
  mx = fmax(mx, dpp_d<0x128>(mx));        // row_ror:8
  const double shf = alive ? exp(lw - mx) : -1.0;
  const double ew = alive ? exp(lw) : 0.0;
  const double vs = alive ? shf : 0.0, ve = alive ? ew : 0.0;
  double cs = vs, ce = ve;
  for (int t = 1; t < n; ++t) {
    const double ps = dpp_d<0x111>(cs);   // row_shr:1: lane t reads lane t - 1
    const double pe = dpp_d<0x111>(ce);
    if (rl == t) {
      cs = ps + vs;
      ce = pe + ve;
    }
  }
  const double ssum = __shfl(cs, rb0 + n - 1, 64);
  const double rb = __shfl(ce, rb0 + n - 1, 64);
  const uint32_t arow = (uint32_t)((__ballot(alive) >> rb0) & 0xffffull);
  const int nk = __popc(arow);
  const double ws = u * ssum;
  const uint32_t crow = (uint32_t)((__ballot(alive && !(cs < ws)) >> rb0) & 0xffffull);
  int sel = -1;
  if (nk > 0) sel = crow ? __ffs((int)crow) - 1 : 31 - __clz((int)arow);
  double W = 0.0;
  const double er_sel = __shfl(er, rb0 + (sel >= 0 ? sel : 0), 64);
  if (nk > 0 && !(rb < 1e-150)) {
    W = rb / (double)n;
    if (!bias) W *= exp(-beta * er_sel);
  } else {
    sel = -1;
  }
  sel_out = sel;
  W_out = W;
}

// Screen thresholds and the cell table of a screening workgroup, in LDS.
template <bool TAB_LDS>
__device__ inline const uint4* stage_screen(const WidomArgs& A, char* dyn, float*& thr)
{
  const int T = A.bc.ntypes;
  thr = reinterpret_cast<float*>(dyn + A.lds_thr);
  for (int i = threadIdx.x; i < A.M * T; i += blockDim.x) {
    const int a = i / T, ty = i % T;
    thr[i] = screen_threshold(A.ff64 + 4 * (ty * T + A.site_type[a]), A.bc.overlap, A.bc.cut_vdw2);
  }
  if (TAB_LDS) {
    uint4* s = reinterpret_cast<uint4*>(dyn + A.lds_tab);
    const int n = A.cell_n[0] * A.cell_n[1] * A.cell_n[2] * A.cell_cap;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s[i] = A.ctab[i];
    return s;
  }
  return A.ctab;
}

// Records of an insertion whose construction failed (W = 0).
__device__ inline void failed_record(const WidomArgs& A, int ins, int fb_placed)
{
  A.rec_w[ins] = 0.0;
  A.rec_e[ins] = 0.0;
  A.rec_f[ins] = fb_placed ? 1 : 0;
  if (A.res) {
    graspa_widom_result r;
    r.rosen = 0.0; r.energy = 0.0; r.vdw = 0.0; r.real = 0.0; r.ewald = 0.0;
    r.first_bead = -1; r.orientation = -1; r.ok = 0; r._pad = 0;
    A.res[ins] = r;
  }
}

// A: first-bead trials, one lane per trial (64 / K insertions per wave step):
// the trial's three draws and its position; the voxel map decides most trials
// with one bit, the others are queued in LDS and screened against the cell
// grid by the whole workgroup, one trial per lane.  Per insertion the
// screened trials go to mask1; an insertion with no trial left is finished
// here (W = 0, most insertions in a dense framework), the others go to the
// live list and their surviving trial sites to the workgroup's sub-list of
// list 1.
template <bool CUBIC, bool TAB_LDS>
__global__ __launch_bounds__(kFbWaves * 64) void widom_first_bead_kernel(const WidomArgs A)
{
  extern __shared__ __align__(16) char dyn[];
  __shared__ unsigned wg_cnt, q_n[2], live_cnt;
  __shared__ unsigned long long wg_st;
  __shared__ uint4 q_p[kFbWaves * 64];             // undecided trial sites (fixed point)
  __shared__ unsigned short q_t[kFbWaves * 64];    // their thread
  __shared__ unsigned char q_hit[kFbWaves * 64];   // screen result by thread
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if (tid == 0) {
    wg_cnt = 0u;
    q_n[0] = 0u;
    q_n[1] = 0u;
    live_cnt = 0u;
    wg_st = 0ull;
  }
  float* thr;
  const uint4* tab = stage_screen<TAB_LDS>(A, dyn, thr);
  __syncthreads();
  const int K = A.K, ipw = 64 / K;
  const int lins = lane / K, trial = lane - lins * K;
  const int i0 = blockIdx.x * A.chunk, i1 = min(A.n_insert, i0 + A.chunk);
  const int per_step = kFbWaves * ipw;
  const int nsteps = i1 > i0 ? (i1 - i0 + per_step - 1) / per_step : 0;
  TrialPt* sub = A.list1 + (size_t)i0 * K;
  const unsigned long long allK = (1ull << K) - 1ull;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const uint32_t* dm = A.dmap[0];
  unsigned long long st = 0;
  for (int step = 0; step < nsteps; ++step) {
    const int ins = i0 + step * per_step + wv * ipw + lins;
    const bool act = lins < ipw && ins < i1;
    double p[3] = {0.0, 0.0, 0.0};
    uint4 P = make_uint4(0u, 0u, 0u, 0u);
    bool hit = false, undecided = false;
    if (act) {
      double u[3];
      draw3(A, (uint64_t)(first_of(A) + ins), 3 * trial, u);
      p[0] = A.bc.cell[0] * u[0];
      p[1] = A.bc.cell[4] * u[1];
      p[2] = A.bc.cell[8] * u[2];
      P = to_fixed4(A.bc, p);
      if (dm) hit = voxel_bit(dm, A.dnv, P);
      // a trial in a certainly free voxel skips the cell screen (it could
      // not flag it)
      undecided = !hit && A.use_cells && !(A.fmap0 && voxel_bit(A.fmap0, A.dnv, P));
    }
    // queue the undecided trials of the workgroup
    unsigned* qn = &q_n[step & 1];
    const unsigned long long um = __ballot(undecided);
    if (um) {
      unsigned b = 0;
      if (lane == 0) b = atomicAdd(qn, (unsigned)__popcll(um));
      b = (unsigned)__builtin_amdgcn_readfirstlane((int)b);
      if (undecided) {
        const unsigned r = b + (unsigned)__popcll(um & lt);
        q_p[r] = P;
        q_t[r] = (unsigned short)tid;
      }
    }
    __syncthreads();
    if (tid == 0) q_n[(step + 1) & 1] = 0u;
    // the queued trials, nine lanes each (lane 9t + r of a wave scans row r
    // of queued trial t's neighbourhood; 7 trials per wave), rows joined by
    // a ballot: a single lane per trial would walk the 27 cells in series
    const unsigned nq = *qn;
    {
      const int tw = lane / 9, tr = lane - 9 * tw;
      constexpr int kPerRound = kFbWaves * 7;
      const int rounds = ((int)nq + kPerRound - 1) / kPerRound;
      for (int rd = 0; rd < rounds; ++rd) {
        const int q = rd * kPerRound + wv * 7 + tw;
        const bool sact = tw < 7 && q < (int)nq;
        const bool h = sact && screen_row<CUBIC>(A, tab, thr, q_p[q], tr, st);
        const unsigned long long hb = __ballot(h);
        if (sact && tr == 0) q_hit[q_t[q]] = ((hb >> (9 * tw)) & 0x1ffull) != 0ull ? 1 : 0;
      }
    }
    __syncthreads();
    if (undecided) hit = q_hit[tid] != 0;
    const unsigned long long hits = __ballot(act && hit);
    const unsigned long long alive = __ballot(act && !hit);
    const bool lead = act && trial == 0;
    const bool live = lead && ((alive >> (lins * K)) & allK) != 0ull;
    if (lead) {
      A.mask1[ins] = (uint32_t)((hits >> (lins * K)) & allK);
      if (!live) failed_record(A, ins, 0);
    }
    const unsigned long long lm = __ballot(live);
    if (lm) {
      unsigned b = 0;
      if (lane == 0) b = atomicAdd(&live_cnt, (unsigned)__popcll(lm));
      b = (unsigned)__builtin_amdgcn_readfirstlane((int)b);
      if (live) A.live[i0 + b + __popcll(lm & lt)] = ins;
    }
    if (alive) {
      unsigned b = 0;
      if (lane == 0) b = atomicAdd(&wg_cnt, (unsigned)__popcll(alive));
      b = (unsigned)__builtin_amdgcn_readfirstlane((int)b);
      if (act && !hit) {
        TrialPt q;
        q.x = p[0]; q.y = p[1]; q.z = p[2];
        q.site = 0;
        q.slot = ins * K + trial;
        q.ins = ins;
        q.bit = trial;
        sub[b + __popcll(alive & lt)] = q;
      }
    }
  }
  if (st) atomicAdd(&wg_st, st);   // every lane's count (each lane screened its own trials)
  __syncthreads();
  if (tid == 0) {
    A.counts1[blockIdx.x] = wg_cnt;
    A.live_counts[blockIdx.x] = live_cnt;
    A.stat_part[blockIdx.x] = wg_st;
  }
}

// a * b rounded on its own: never fused with the operation that uses it (the
// cell fractions of the pair kernel must round alike whether they are staged
// in LDS or computed as the framework is read).
__device__ inline double mul_unfused(double a, double b)
{
#pragma clang fp contract(off)
  return a * b;
}

// Atoms of the pair kernel's LDS framework copy: whole steps of 128 (two per lane).
__host__ __device__ constexpr int eval_fw_pad(int n) { return (n + 127) & ~127; }
// Slot of atom i's type in that copy: the types of atoms 128 s + l and
// 128 s + 64 + l sit side by side (one 8-byte read per lane and step).
__device__ inline int fw_type_slot(int i) { return ((i >> 7) << 7) + ((i & 63) << 1) + ((i >> 6) & 1); }
// Slot of atom i's (x, y) in that copy (double2 units), its (z, q) 64 slots
// further: a step of 128 atoms is four blocks of 64 double2, (x, y) and
// (z, q) of atoms 128 s + l, then of atoms 128 s + 64 + l, so the 64 lanes'
// 16-byte reads of a block are contiguous (no two lanes of a pass share a
// bank, where 32-byte records put lanes l and l + 8 on the same banks).
__device__ inline int fw_xy_slot(int i) { return ((i >> 7) << 8) + (((i >> 6) & 1) << 7) + (i & 63); }

// B / D: full energies of the trial sites of a work list (VDW_Coulomb.cu:
// 1183-1352 for one trial site): wave sums of LJ and real-space Coulomb and
// the exact overlap flag; one wave per site, 64 framework atoms per step.
// The framework (x, y, z, q and type) is staged in LDS when it fits.
template <bool MIXED, bool CUBIC, bool FW_LDS, int W>
__global__ __launch_bounds__(W * 64) __attribute__((amdgpu_waves_per_eu(W == 16 ? 8 : 1))) void widom_eval_kernel(const WidomArgs A, const TrialPt* list,
                                                                   const unsigned* counts, int groups, int chunk,
                                                                   int per_ins, double2* energy, uint32_t* mask,
                                                                   unsigned long long* stat_slot)
{
  extern __shared__ __align__(16) char dyn[];
  __shared__ unsigned long long wg_eval;
  if (threadIdx.x == 0) wg_eval = 0ull;
  // exclusive prefix sums of the sub-list lengths (groups sub-lists of
  // chunk * per_ins entries each)
  __shared__ unsigned pre[kMaxScreenGroups + 1];
  const int G = groups;
  for (int i = threadIdx.x; i < G; i += blockDim.x) pre[i + 1] = counts[i];
  if (threadIdx.x == 0) pre[0] = 0u;
  __syncthreads();
  for (int h = 1; h < G; h <<= 1) {
    unsigned add[(kMaxScreenGroups + W * 64 - 1) / (W * 64)];
    for (int k = 0, i = threadIdx.x; i < G; ++k, i += blockDim.x) add[k] = (i >= h) ? pre[i + 1 - h] : 0u;
    __syncthreads();
    for (int k = 0, i = threadIdx.x; i < G; ++k, i += blockDim.x) pre[i + 1] += add[k];
    __syncthreads();
  }
  const BoxConst& bc = A.bc;
  const int T = bc.ntypes, n = A.n_fw;
  double4* ffl = reinterpret_cast<double4*>(dyn + A.lds_ffl);
  for (int i = threadIdx.x; i < A.M * T; i += blockDim.x) {
    const int a = i / T, ty = i % T;
    const double* f = A.ff64 + 4 * (ty * T + A.site_type[a]);
    if (MIXED) {
      // fast form: {c12, c6, shift} = {4 eps sigma^12, 4 eps sigma^6, shift}
      const double s6 = f[1] * f[1] * f[1], c6 = f[0] * s6;
      ffl[i] = make_double4(c6 * s6, c6, f[2], 0.0);
    } else {
      ffl[i] = make_double4(f[0], f[1], f[2], 0.0);
    }
  }
  // the table sits at LDS offset 0 (lds_coul == 0): a row's address is one
  // multiply of its index
  double2* ct = reinterpret_cast<double2*>(dyn);
  if (MIXED)
    for (int i = threadIdx.x; i < A.coul_n / 2; i += blockDim.x)
      ct[(i / (kCoulTabStride / 2)) * kCoulLdsRow + i % (kCoulTabStride / 2)] = reinterpret_cast<const double2*>(A.coul)[i];
  // the LDS copy is padded to whole steps of two atoms per lane with atoms at
  // NaN positions: every comparison with their r^2 is false, so they add
  // nothing, flag nothing and the steps need no bound check
  const int npad = eval_fw_pad(n);
  double2* fw2 = reinterpret_cast<double2*>(dyn + A.lds_fw);
  int* fwt = reinterpret_cast<int*>(dyn + A.lds_fw + (size_t)npad * sizeof(double4));
  if (FW_LDS) {
    const double qnan = __builtin_nan("");
    // fast form in an equal-edged cube: x, y, z in cell fractions
    const double sc = (MIXED && A.iso) ? bc.inv[0] : 1.0;
    for (int i = threadIdx.x; i < npad; i += blockDim.x) {
      // (products rounded on their own, as the path that reads the framework
      // from memory rounds them)
      const int xy = fw_xy_slot(i);
      fw2[xy] = i < n ? make_double2(mul_unfused(A.fw[i], sc), mul_unfused(A.fw[n + i], sc)) : make_double2(qnan, qnan);
      fw2[xy + 64] = i < n ? make_double2(mul_unfused(A.fw[2 * n + i], sc), A.fw[3 * n + i]) : make_double2(qnan, 0.0);
      fwt[fw_type_slot(i)] = i < n ? A.fw_type[i] : 0;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int npts = (int)pre[G];
  unsigned long long n_eval = 0;
  // which wave evaluates which point does not change any result.  A long
  // list is cut into a contiguous range per wave (one binary search for its
  // first point's sub-list, then the sub-list index only moves forward); a
  // short one (under four points per wave) is dealt out point by point (cut
  // into ranges, the orientation list of a batch took 36 us instead of 29)
  const int nwave = gridDim.x * W, wid = blockIdx.x * W + wv;
  const int per_wave = (npts + nwave - 1) / nwave;
  const bool ranges = per_wave >= 4;
  const int p_begin = ranges ? wid * per_wave : wid, p_end = ranges ? min(npts, p_begin + per_wave) : npts;
  const int p_step = ranges ? 1 : nwave;
  int lo = 0;
  for (int pi = p_begin; pi < p_end; pi += p_step) {
    if (!ranges || pi == p_begin) {
      // sub-list g holding point pi: the last g with pre[g] <= pi
      int hi = G - 1;
      lo = 0;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)pre[mid] <= pi) lo = mid; else hi = mid - 1;
      }
    }
    while (lo + 1 < G && (int)pre[lo + 1] <= pi) ++lo;
    const TrialPt P = list[(size_t)lo * chunk * per_ins + (pi - (int)pre[lo])];
    const double px = P.x, py = P.y, pz = P.z;
    const double sq = A.site_q[P.site];
    const bool coul = bc.use_charges && sq != 0.0;
    const double4* fr = ffl + P.site * T;
    double v = 0.0, r = 0.0;
    bool o = false;
    if (MIXED) {
      // two atoms per lane per step (independent chains).  The LJ term and
      // the real-space term sit behind lane-divergent cutoff branches (one
      // branch when the two cutoffs are equal): the force-field row and the
      // table row are read only by lanes inside them, and a step whose atoms
      // are all beyond the cutoffs for every lane skips both terms.  r sums
      // q_atom g(r^2) (one FMA); the prefactor and the site's charge multiply
      // the wave's sum once
      double v1 = 0.0, r1 = 0.0;
      bool o1 = false;
      // squared distance to the site: in an equal-edged cube from cell
      // fractions (13 instructions: per axis a difference, its nearest
      // integer removed; then L^2 times the sum), else from Cartesian
      // coordinates through pbc_rint (15)
      const double Lc = bc.cell[0], L2 = Lc * Lc, iL = bc.inv[0];
      const double fx = mul_unfused(px, iL), fy = mul_unfused(py, iL), fz = mul_unfused(pz, iL);
      auto run = [&](auto iso_c, auto same_c, auto general_c) {
        constexpr bool ISO = decltype(iso_c)::value;      // distances from cell fractions
        constexpr bool SAME = decltype(same_c)::value;    // equal LJ and Coulomb cutoffs
        // general form: the shift is read and r^2 < 0.01 tested; else every
        // pair closer than 0.1 A has an LJ term above OverlapCriteria (or NaN
        // at r = 0), so !(e <= overlap) alone flags it as the reference does
        constexpr bool GENERAL = decltype(general_c)::value;
        auto lj = [&](double rr, int ty, double& v_, bool& o_) {
          const double4* f = fr + ty;
          const double2 cc = *reinterpret_cast<const double2*>(f);
          const double e = lj_mixed_c(rr, cc.x, cc.y, GENERAL ? f->z : 0.0);
          v_ += e;
          o_ = o_ || !(e <= bc.overlap) || (GENERAL && rr < 0.01);
        };
        auto pair_fast = [&](double rr, double aq, int ty, double& v_, double& r_, bool& o_) {
          if constexpr (SAME) {
            if (rr < bc.cut_vdw2) {
              lj(rr, ty, v_, o_);
              if (coul) r_ = fma(aq, coul_table(ct, rr), r_);
            }
          } else {
            if (rr < bc.cut_vdw2) lj(rr, ty, v_, o_);
            if (coul && rr < bc.cut_coul2) r_ = fma(aq, coul_table(ct, rr), r_);
          }
        };
        // (the sums as explicit FMAs: both loops below round them alike, so
        // every launch shape gives the same sums)
        auto dist2 = [&](double ax, double ay, double az) {
          if constexpr (ISO) {
            double dx = ax - fx, dy = ay - fy, dz = az - fz;
            dx -= rint(dx); dy -= rint(dy); dz -= rint(dz);
            return L2 * fma(dz, dz, fma(dy, dy, dx * dx));
          } else {
            double dx = ax - px, dy = ay - py, dz = az - pz;
            pbc_rint<CUBIC>(dx, dy, dz, bc.cell, bc.inv);
            return fma(dz, dz, fma(dy, dy, dx * dx));
          }
        };
        if (FW_LDS) {
          // (a wave-uniform step count: the loop test is scalar)
          for (int it = 0, nit = npad >> 7; it < nit; ++it) {
            const int j = (it << 7) + lane;
            const double2* st = fw2 + (it << 8) + lane;
            const double2 xy0 = st[0], zq0 = st[64], xy1 = st[128], zq1 = st[192];
            const double4 a0 = make_double4(xy0.x, xy0.y, zq0.x, zq0.y), a1 = make_double4(xy1.x, xy1.y, zq1.x, zq1.y);
            // types of atoms j and j + 64 in one 8-byte read (kept whole rather
            // than split into two reads, one of them under the cutoff branch)
            const unsigned long long tt =
                __builtin_nontemporal_load(reinterpret_cast<const unsigned long long*>(fwt) + (j - 64 * it));
            const int ty0 = (int)(unsigned)tt, ty1 = (int)(unsigned)(tt >> 32);
            const double rr0 = dist2(a0.x, a0.y, a0.z);
            const double rr1 = dist2(a1.x, a1.y, a1.z);
            pair_fast(rr0, a0.w, ty0, v, r, o);
            pair_fast(rr1, a1.w, ty1, v1, r1, o1);
          }
        } else {
          const double sc = ISO ? iL : 1.0;
          for (int j = lane; j < n; j += 128) {
            const int k = j + 64;
            const bool has1 = k < n;
            const double ax0 = mul_unfused(A.fw[j], sc), ay0 = mul_unfused(A.fw[n + j], sc);
            const double az0 = mul_unfused(A.fw[2 * n + j], sc);
            const double aq0 = A.fw[3 * n + j];
            const int ty0 = A.fw_type[j];
            double ax1 = 0.0, ay1 = 0.0, az1 = 0.0, aq1 = 0.0;
            int ty1 = 0;
            if (has1) {
              ax1 = mul_unfused(A.fw[k], sc); ay1 = mul_unfused(A.fw[n + k], sc); az1 = mul_unfused(A.fw[2 * n + k], sc);
              aq1 = A.fw[3 * n + k];
              ty1 = A.fw_type[k];
            }
            const double rr0 = dist2(ax0, ay0, az0);
            const double rr1 = has1 ? dist2(ax1, ay1, az1) : 1e300;   // 1e300: contributes nothing
            pair_fast(rr0, aq0, ty0, v, r, o);
            pair_fast(rr1, aq1, ty1, v1, r1, o1);
          }
        }
      };
      auto run_same = [&](auto iso_c, auto general_c) {
        if (bc.cut_vdw2 == bc.cut_coul2) run(iso_c, std::true_type{}, general_c);
        else run(iso_c, std::false_type{}, general_c);
      };
      auto run_general = [&](auto iso_c) {
        if (A.lj_general) run_same(iso_c, std::true_type{});
        else run_same(iso_c, std::false_type{});
      };
      if (CUBIC && A.iso) run_general(std::true_type{});
      else run_general(std::false_type{});
      v += v1;
      r += r1;
      o = o || o1;
    } else {
      // plain FP64 form: the reference's expressions, one atom at a time
      for (int j = lane; j < n; j += 64) {
        double ax, ay, az, aq;
        int ty;
        if (FW_LDS) {
          const double2 xy = fw2[fw_xy_slot(j)], zq = fw2[fw_xy_slot(j) + 64];
          ax = xy.x; ay = xy.y; az = zq.x; aq = zq.y; ty = fwt[fw_type_slot(j)];
        } else {
          ax = A.fw[j]; ay = A.fw[n + j]; az = A.fw[2 * n + j]; aq = A.fw[3 * n + j]; ty = A.fw_type[j];
        }
        double dx = ax - px, dy = ay - py, dz = az - pz;
        pbc_fp64(dx, dy, dz, bc.cell, bc.inv, CUBIC ? 1 : 0);
        const double rr = dx * dx + dy * dy + dz * dz;
        if (rr < bc.cut_vdw2) {
          const double4 ff = fr[ty];
          const double e = lj_fp64(rr, ff.x, ff.y, ff.z);
          o = o || (e > bc.overlap) || (rr < 0.01);
          v += e;
        }
        if (coul && rr < bc.cut_coul2) {
          const double rd = sqrt(rr);
          r += bc.prefactor * 1.0 * (aq * sq * erfc(bc.alpha * rd)) / rd;
        }
      }
    }
    // guest molecules (guest-guest), from memory
    unsigned long long npair = (unsigned long long)n;
    for (int c = 1; c < A.gv.n_comp; ++c) {
      const int ng = A.gv.seg[c].n_atoms, base = A.gv.seg[c].base;
      npair += (unsigned long long)ng;
      for (int jj = lane; jj < ng; jj += 64) {
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
          const double e = MIXED ? lj_mixed_c(rr, ff.x, ff.y, ff.z) : lj_fp64(rr, ff.x, ff.y, ff.z);
          o = o || (e > bc.overlap) || (rr < 0.01);
          v += e;
        }
        if (coul && rr < bc.cut_coul2) {
          if (MIXED) {
            r = fma(gq, coul_table(ct, rr), r);    // scaled with the framework sum below
          } else {
            const double rd = sqrt(rr);
            r += bc.prefactor * 1.0 * (gq * sq * erfc(bc.alpha * rd)) / rd;
          }
        }
      }
    }
    n_eval += npair;
    const double sv = wave_sum(v);
    const double sr = coul ? (MIXED ? (bc.prefactor * sq) * wave_sum(r) : wave_sum(r)) : 0.0;
    const bool ov = __any(o) != 0;
    if (lane == 0) {
      energy[P.slot] = make_double2(sv, sr);
      if (ov) atomicOr(mask + P.ins, 1u << P.bit);
    }
  }
  if (lane == 0 && n_eval) atomicAdd(&wg_eval, n_eval);
  __syncthreads();
  if (threadIdx.x == 0) stat_slot[blockIdx.x] = wg_eval;
}

// C: first-bead selection; orientations about the selected bead and their
// overlap screen.  Workgroup g owns insertions [g*chunk, (g+1)*chunk).
// C: one wave per live insertion (workgroup g takes the live sub-list of
// first-bead workgroup g): first-bead selection from the evaluated trials,
// the O orientations of the template about the selected bead (one lane per
// orientation site), their overlap screen, survivors appended to the
// workgroup's sub-list of list 2.
template <bool CUBIC>
__global__ __launch_bounds__(kInsWaves * 64) void widom_chain_kernel(const WidomArgs A)
{
  extern __shared__ __align__(16) char dyn[];
  __shared__ unsigned wg_cnt, wg_pl;
  __shared__ unsigned long long wg_st;
  const int T = A.bc.ntypes;
  float* thr = reinterpret_cast<float*>(dyn);
  for (int i = threadIdx.x; i < A.M * T; i += blockDim.x) {
    const int a = i / T, ty = i % T;
    thr[i] = screen_threshold(A.ff64 + 4 * (ty * T + A.site_type[a]), A.bc.overlap, A.bc.cut_vdw2);
  }
  if (threadIdx.x == 0) {
    wg_cnt = 0u;
    wg_pl = 0u;
    wg_st = 0ull;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int K = A.K, O = A.O, chain = A.chain;
  const int i0 = blockIdx.x * A.chunk;
  const unsigned n_live = A.live_counts[blockIdx.x];
  TrialPt* sub = A.list2 + (size_t)i0 * O * chain;
  const unsigned long long lt = (1ull << lane) - 1ull;
  unsigned long long st = 0;
  int* placed = A.placed + i0;
  // the placed insertions of this wave step (ballot bits `pm`, insertion in
  // the lane of each bit) join the workgroup's placed list for the finish
  // kernel; a live insertion whose first bead was not placed gets its W = 0
  // record here
  auto to_finish = [&](const unsigned long long pm, const bool on_lane, const int ins) {
    if (!pm) return;
    unsigned b = 0;
    if (lane == 0) b = atomicAdd(&wg_pl, (unsigned)__popcll(pm));
    b = (unsigned)__builtin_amdgcn_readfirstlane((int)b);
    if (on_lane) placed[b + __popcll(pm & lt)] = ins;
  };
  // orientations of a live insertion whose first bead was placed (the whole
  // wave; lane o * chain + a is orientation o's chain atom a)
  auto orient = [&](const int ins, const uint64_t idx, const double fb0, const double fb1, const double fb2) {
    const int o = lane / chain, a = lane - o * chain;
    const bool act = lane < O * chain;
    double p[3] = {0.0, 0.0, 0.0};
    bool hit = false;
    if (act) {
      double u[3];
      draw3(A, idx, 3 * K + 1 + 3 * o, u);
      double vec[3] = {A.tvec[a][0], A.tvec[a][1], A.tvec[a][2]};
      quat_rotate(vec, u[0], u[1], u[2]);
      p[0] = fb0 + vec[0]; p[1] = fb1 + vec[1]; p[2] = fb2 + vec[2];
      const uint4 P = to_fixed4(A.bc, p);
      if (A.dmap[1 + a]) hit = voxel_bit(A.dmap[1 + a], A.dnv, P);
      if (!hit && A.screen_chain) hit = screen_lane<CUBIC>(A, A.ctab, thr + (1 + a) * T, P, st);
    }
    const unsigned long long hits = __ballot(act && hit);
    const unsigned long long am = (1ull << chain) - 1ull;
    uint32_t dead2 = 0u;
    for (int oo = 0; oo < O; ++oo)
      if ((hits >> (oo * chain)) & am) dead2 |= 1u << oo;
    if (lane == 0) A.mask2[ins] = dead2;
    const bool alive = act && !((dead2 >> o) & 1u);
    const unsigned long long al = __ballot(alive);
    if (al) {
      unsigned b = 0;
      if (lane == 0) b = atomicAdd(&wg_cnt, (unsigned)__popcll(al));
      b = (unsigned)__builtin_amdgcn_readfirstlane((int)b);
      if (alive) {
        TrialPt q;
        q.x = p[0]; q.y = p[1]; q.z = p[2];
        q.site = 1 + a;
        q.slot = (ins * O + o) * chain + a;
        q.ins = ins;
        q.bit = o;
        sub[b + __popcll(al & lt)] = q;
      }
    }
  };
  if (K <= 16) {
    // first-bead selections of four live insertions per wave, one per
    // 16-lane row; then the orientations of those whose bead was placed
    const int row = lane >> 4, rl = lane & 15;
    for (unsigned b0 = wv * 4; b0 < n_live; b0 += kInsWaves * 4) {
      const unsigned li = b0 + row;
      const bool has = li < n_live;
      const int ins = has ? A.live[i0 + li] : 0;
      const uint64_t idx = (uint64_t)(first_of(A) + ins);
      const uint32_t dead = has ? A.mask1[ins] : ~0u;
      const bool alive = has && rl < K && !((dead >> rl) & 1u);
      double ev = 0.0, er = 0.0;
      if (alive) {
        const double2 e = A.e1[(size_t)ins * K + rl];
        ev = e.x;
        er = e.y;
      }
      int sel;
      double Wfb;
      row_select(ev, er, alive, K, draw(A, idx, 3 * K), A.beta, A.bias, sel, Wfb);
      const bool fb_ok = has && sel >= 0 && Wfb > 1e-150;
      const int sl = row * 16 + (fb_ok ? sel : 0);
      InsState S;
      S.sel = fb_ok ? sel : -1;
      S.ok = fb_ok ? 1 : 0;
      S.wfb = Wfb;
      const double ev_s = __shfl(ev, sl, 64), er_s = __shfl(er, sl, 64);
      S.ev = fb_ok ? ev_s : 0.0;
      S.er = fb_ok ? er_s : 0.0;
      S.fb[0] = S.fb[1] = S.fb[2] = 0.0;
      if (fb_ok) {
        double u[3];
        draw3(A, idx, 3 * sel, u);
        S.fb[0] = A.bc.cell[0] * u[0];
        S.fb[1] = A.bc.cell[4] * u[1];
        S.fb[2] = A.bc.cell[8] * u[2];
      }
      if (has && rl == 0) {
        A.state[ins] = S;
        A.mask2[ins] = 0u;
        if (!fb_ok) failed_record(A, ins, 0);
      }
      const unsigned long long okm = __ballot(fb_ok && rl == 0);
      to_finish(okm, fb_ok && rl == 0, ins);
      if (chain > 0) {
        for (int r = 0; r < 4; ++r) {
          if (!((okm >> (16 * r)) & 1ull)) continue;
          const int ins_r = __builtin_amdgcn_readlane(ins, 16 * r);
          orient(ins_r, (uint64_t)(first_of(A) + ins_r), readlane_d(S.fb[0], 16 * r), readlane_d(S.fb[1], 16 * r),
                 readlane_d(S.fb[2], 16 * r));
        }
      }
    }
  } else {
    for (unsigned li = wv; li < n_live; li += kInsWaves) {
      const int ins = A.live[i0 + li];
      const uint64_t idx = (uint64_t)(first_of(A) + ins);
      const uint32_t dead = A.mask1[ins];
      double ev = 0.0, er = 0.0;
      if (lane < K && !((dead >> lane) & 1u)) {
        const double2 e = A.e1[(size_t)ins * K + lane];
        ev = e.x;
        er = e.y;
      }
      int sel;
      double Wfb;
      wave_select(ev, er, dead, K, draw(A, idx, 3 * K), A.beta, A.bias, sel, Wfb);
      const bool fb_ok = sel >= 0 && Wfb > 1e-150;
      InsState S;
      S.sel = fb_ok ? sel : -1;
      S.ok = fb_ok ? 1 : 0;
      S.wfb = Wfb;
      S.ev = fb_ok ? readlane_d(ev, sel) : 0.0;
      S.er = fb_ok ? readlane_d(er, sel) : 0.0;
      S.fb[0] = S.fb[1] = S.fb[2] = 0.0;
      if (fb_ok) {
        double u[3];
        draw3(A, idx, 3 * sel, u);
        S.fb[0] = A.bc.cell[0] * u[0];
        S.fb[1] = A.bc.cell[4] * u[1];
        S.fb[2] = A.bc.cell[8] * u[2];
      }
      if (lane == 0) {
        A.state[ins] = S;
        A.mask2[ins] = 0u;
        if (!fb_ok) failed_record(A, ins, 0);
      }
      to_finish(fb_ok ? 1ull : 0ull, fb_ok && lane == 0, ins);
      if (fb_ok && chain > 0) orient(ins, idx, S.fb[0], S.fb[1], S.fb[2]);
    }
  }
  if (st) atomicAdd(&wg_st, st);   // every lane's count (each lane screened its own trials)
  __syncthreads();
  if (threadIdx.x == 0) {
    A.counts2[blockIdx.x] = wg_cnt;
    A.placed_counts[blockIdx.x] = wg_pl;
    A.stat_part[kMaxScreenGroups + blockIdx.x] = wg_st;
  }
}

// E: orientation selection, Fourier-space energy of the selected molecule
// (Fourier_Ewald_Diff, Ewald_Energy_Functions.h:280-397, with Oldsize = 0)
// and the record: W = W_fb W_ch exp(-beta (dE_ewald + dE_tail)), for the
// insertions whose first bead was placed (the others have their records from
// the first-bead or the chain kernel).  gridDim.x / groups workgroups share
// the placed list of screening workgroup g, one wave per insertion.
__global__ __launch_bounds__(kFinWaves * 64) void widom_finish_kernel(const WidomArgs A)
{
  extern __shared__ __align__(16) char dyn[];
  const BoxConst& bc = A.bc;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int K = A.K, O = A.O, chain = A.chain, M = A.M, KS = A.kstride;
  const size_t wave_bytes = (size_t)M * 3 * KS * sizeof(double2) + kMaxMol * 3 * sizeof(double);
  double2* eik = reinterpret_cast<double2*>(dyn + (size_t)wv * wave_bytes);
  double (*mol)[3] = reinterpret_cast<double (*)[3]>(dyn + (size_t)wv * wave_bytes + (size_t)M * 3 * KS * sizeof(double2));
  const int per = gridDim.x / A.groups, g = blockIdx.x / per, part = blockIdx.x - g * per;
  const unsigned n_pl = A.placed_counts[g];
  const int* placed = A.placed + (size_t)g * A.chunk;
  for (unsigned k = (unsigned)(part * kFinWaves + wv); k < n_pl; k += (unsigned)(per * kFinWaves)) {
    const int ins = placed[k];
    const uint64_t idx = (uint64_t)(first_of(A) + ins);
    const InsState S = A.state[ins];
    bool ok = S.ok != 0;
    double Wtot = S.wfb, e_v = S.ev, e_r = S.er;
    int orient = 0;
    // selects, not an index into the loaded struct (which would place it in scratch)
    if (lane < 3) mol[0][lane] = lane == 0 ? S.fb[0] : (lane == 1 ? S.fb[1] : S.fb[2]);
    if (ok && chain > 0) {
      const uint32_t dead2 = A.mask2[ins];
      double cv = 0.0, cr = 0.0;
      if (lane < O && !((dead2 >> lane) & 1u)) {
        for (int a = 0; a < chain; ++a) {
          const double2 e = A.e2[((size_t)ins * O + lane) * chain + a];
          cv += e.x;
          cr += e.y;
        }
      }
      int so;
      double Wch;
      wave_select(cv, cr, dead2, O, draw(A, idx, 3 * K + 3 * O + 1), A.beta, A.bias, so, Wch);
      if (so < 0 || !(S.wfb * Wch > 1e-150)) {
        ok = false;
      } else {
        Wtot = S.wfb * Wch;
        orient = so;
        e_v += readlane_d(cv, so);
        e_r += readlane_d(cr, so);
        if (lane < chain) {
          double p[3];
          orient_pos(A, idx, so, lane, S.fb, p);
          mol[1 + lane][0] = p[0]; mol[1 + lane][1] = p[1]; mol[1 + lane][2] = p[2];
        }
      }
    }
    wave_sync();
    double ew = 0.0;
    if (ok && A.charged) {
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
    }
    if (lane == 0) {
      double Wf = 0.0, etot = 0.0;
      if (ok) {
        Wf = Wtot * exp(-A.beta * ew) * exp(-A.beta * A.tail_de);
        etot = e_v + e_r + ew + A.tail_de;
      }
      A.rec_w[ins] = Wf;
      A.rec_e[ins] = etot;
      A.rec_f[ins] = (S.ok ? 1 : 0) | (ok ? 2 : 0);
      if (A.res) {
        graspa_widom_result r;
        r.rosen = Wf;
        r.energy = etot;
        r.vdw = ok ? e_v : 0.0;
        r.real = ok ? e_r : 0.0;
        r.ewald = ok ? ew : 0.0;
        r.first_bead = ok ? S.sel : -1;
        r.orientation = ok ? orient : -1;
        r.ok = ok ? 1 : 0;
        r._pad = 0;
        A.res[ins] = r;
      }
    }
    wave_sync();
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

// The batch's sums of the reduction partials, added to the running totals,
// and its work counts, sums of per-workgroup partials (integers): screened
// pairs of the first-bead and chain workgroups ([0, n_screen) and
// [kMaxScreenGroups, kMaxScreenGroups + n_screen)), then evaluated pairs of
// the two pair-energy launches (kMaxEvalGroups slots each).
__global__ __launch_bounds__(kRedBlock) void widom_final_kernel(const double* part, int nblk, double* out,
                                                                double* tot, const unsigned long long* wpart,
                                                                int n_screen, int n_eval1, int n_eval2,
                                                                unsigned long long* stat, unsigned long long* stat_tot)
{
  __shared__ double s[kSums][kRedBlock];
  __shared__ unsigned long long ws[2][kRedBlock];
  {
    // (integer sums: exact in any order)
    unsigned long long a = 0, e = 0;
    for (int i = threadIdx.x; i < n_screen; i += kRedBlock) a += wpart[i] + wpart[kMaxScreenGroups + i];
    for (int i = threadIdx.x; i < n_eval1; i += kRedBlock) e += wpart[2 * kMaxScreenGroups + i];
    for (int i = threadIdx.x; i < n_eval2; i += kRedBlock) e += wpart[2 * kMaxScreenGroups + kMaxEvalGroups + i];
    ws[0][threadIdx.x] = a;
    ws[1][threadIdx.x] = e;
  }
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
  if (threadIdx.x < kSums) {
    const double v = s[threadIdx.x][0];
    out[threadIdx.x] = v;
    tot[threadIdx.x] += v;   // running totals: batches of a box run in order on its stream
  }
  if (threadIdx.x == 0) {
    unsigned long long x = 0, y = 0;
    for (int i = 0; i < kRedBlock; ++i) { x += ws[0][i]; y += ws[1][i]; }
    stat[0] = x;
    stat[1] = y;
    stat_tot[0] += x;
    stat_tot[1] += y;
  }
}

static size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// Resident workgroups per CU for a kernel, workgroup size and dynamic LDS;
// 0 if it cannot run.
template <typename Kern>
static int occupancy(graspa_box* b, Kern kern, int threads, size_t lds)
{
  if (lds > (size_t)b->max_lds) return 0;
  if (lds > 65536) {
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipGetLastError();
  }
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kern, threads, lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return blocks;
}

// Launches a per-insertion kernel on a grid of resident workgroups.
template <typename Kern>
static int launch_ins(graspa_box* b, Kern kern, const WidomArgs& A, size_t lds, const char* what, int waves,
                      int fixed_grid = 0)
{
  const int occ = occupancy(b, kern, waves * 64, lds);
  if (occ < 1) return fail(GRASPA_ERR_UNSUPPORTED, std::string("widom: no launch shape fits the LDS (") + what + ")");
  const long long need = ((long long)A.n_insert + waves - 1) / waves;
  const long long grid = fixed_grid > 0 ? fixed_grid : std::max(1LL, std::min(need, (long long)occ * b->n_cu));
  hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(waves * 64), lds, b->stream, A);
  return hip_check(hipGetLastError(), what);
}

// Per-box buffers of the pipeline, grown on demand.
static int ensure_pipeline(graspa_box* b, int n_insert, int K, int OC)
{
  const size_t n = (size_t)n_insert;
  const size_t need_e1 = n * K, need_e2 = n * std::max(OC, 1);
  if (n <= b->wp_n && need_e1 <= b->wp_e1 && need_e2 <= b->wp_e2) return GRASPA_OK;
  const size_t cn = std::max(n, b->wp_n), c1 = std::max(need_e1, b->wp_e1), c2 = std::max(need_e2, b->wp_e2);
  hipStreamSynchronize(b->stream);
  box_free(b, b->d_wmask1); box_free(b, b->d_wmask2); box_free(b, b->d_we1); box_free(b, b->d_we2);
  box_free(b, b->d_wlist1); box_free(b, b->d_wlist2); box_free(b, b->d_wstate); box_free(b, b->d_wlive);
  b->d_wmask1 = b->d_wmask2 = nullptr;
  b->d_we1 = b->d_we2 = b->d_wlist1 = b->d_wlist2 = b->d_wstate = b->d_wlive = nullptr;
  b->wp_n = b->wp_e1 = b->wp_e2 = 0;
  int rc = hip_check(box_malloc(b, &b->d_wmask1, cn * sizeof(uint32_t)), "hipMalloc widom pipeline");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_wmask2, cn * sizeof(uint32_t)), "hipMalloc widom pipeline");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_wstate, cn * sizeof(InsState)), "hipMalloc widom pipeline");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_wlive, cn * sizeof(int)), "hipMalloc widom pipeline");
  if (!rc && !b->d_wstat)
    rc = hip_check(box_malloc(b, &b->d_wstat, (2 * kMaxScreenGroups + 2 * kMaxEvalGroups) * sizeof(unsigned long long)),
                   "hipMalloc widom pipeline");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_we1, c1 * sizeof(double2)), "hipMalloc widom pipeline");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_wlist1, c1 * sizeof(TrialPt)), "hipMalloc widom pipeline");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_we2, c2 * sizeof(double2)), "hipMalloc widom pipeline");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_wlist2, c2 * sizeof(TrialPt)), "hipMalloc widom pipeline");
  if (!rc && !b->d_wcounts)
    rc = hip_check(box_malloc(b, &b->d_wcounts, 3 * kMaxScreenGroups * sizeof(unsigned)), "hipMalloc widom pipeline");
  if (!rc) {
    b->wp_n = cn;
    b->wp_e1 = c1;
    b->wp_e2 = c2;
  }
  return rc;
}

using EvalKern = void (*)(const WidomArgs, const TrialPt*, const unsigned*, int, int, int, double2*, uint32_t*,
                          unsigned long long*);

template <bool MIXED, bool CUBIC, bool FW_LDS>
static EvalKern eval_kernel_w(int waves)
{
  switch (waves) {
    case 4: return widom_eval_kernel<MIXED, CUBIC, FW_LDS, 4>;
    case 8: return widom_eval_kernel<MIXED, CUBIC, FW_LDS, 8>;
    case 16: return widom_eval_kernel<MIXED, CUBIC, FW_LDS, 16>;
    default: return widom_eval_kernel<MIXED, CUBIC, FW_LDS, 12>;
  }
}

// The pair-energy kernel of a launch shape for the box's precision form and cell.
static EvalKern eval_kernel_for(const graspa_box* b, int waves, bool fw_lds)
{
  const bool mixed = b->precision == GRASPA_MIXED, cubic = b->bc.cubic != 0;
  if (mixed && cubic) return fw_lds ? eval_kernel_w<true, true, true>(waves) : eval_kernel_w<true, true, false>(waves);
  if (mixed) return fw_lds ? eval_kernel_w<true, false, true>(waves) : eval_kernel_w<true, false, false>(waves);
  if (cubic) return fw_lds ? eval_kernel_w<false, true, true>(waves) : eval_kernel_w<false, true, false>(waves);
  return fw_lds ? eval_kernel_w<false, false, true>(waves) : eval_kernel_w<false, false, false>(waves);
}

// LDS of the pair-energy kernel before the framework: FF rows of the species'
// sites, then the real-space table (fast form).
static size_t eval_lds_coul_bytes(const graspa_box* b)
{
  return b->precision == GRASPA_MIXED ? align16((size_t)coul_lds_doubles(b->coul_n) * sizeof(double)) : 0;
}
static size_t eval_lds_base(const graspa_box* b, int M)
{
  return align16(eval_lds_coul_bytes(b) + (size_t)M * b->bc.ntypes * sizeof(double4));
}

static size_t eval_lds(const graspa_box* b, size_t base, bool fw_lds)
{
  return base + (fw_lds ? align16((size_t)eval_fw_pad(b->n_fw) * (sizeof(double4) + sizeof(int))) : 0);
}

// Grid of one pair-energy launch: per_cu workgroups per CU (0: as many as
// are resident), capped by the occupancy and the work-count slots; 0 if the
// shape does not fit.
static int eval_grid(graspa_box* b, EvalKern kern, int waves, size_t lds, int per_cu)
{
  const int occ = occupancy(b, kern, waves * 64, lds);
  if (occ < 1) return 0;
  const int wpc = per_cu > 0 ? std::min(per_cu, occ) : occ;
  return std::min(kMaxEvalGroups, wpc * b->n_cu);
}

// One pair-energy launch in the box's launch shape.  A shape that stages the
// framework in LDS falls back to reading it from L2 when it does not fit.
static int launch_eval(graspa_box* b, const WidomArgs& A, const TrialPt* list, const unsigned* counts, int groups,
                       int chunk, int per_ins, double2* energy, uint32_t* mask, unsigned long long* stat_slot,
                       int* grid_out, bool timed)
{
  const int waves = b->ev_waves;
  for (int pass = b->ev_fw_lds ? 0 : 1; pass < 2; ++pass) {
    const bool fw_lds = pass == 0;
    const size_t lds = eval_lds(b, (size_t)A.lds_fw, fw_lds);
    const EvalKern kern = eval_kernel_for(b, waves, fw_lds);
    const int grid = eval_grid(b, kern, waves, lds, b->ev_per_cu);
    if (grid < 1) continue;
    *grid_out = grid;
    if (timed) timing_begin(b, "widom");
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(waves * 64), lds, b->stream, A, list, counts, groups, chunk,
                       per_ins, energy, mask, stat_slot);
    if (timed) timing_end(b, "widom");
    return hip_check(hipGetLastError(), "widom_eval_kernel launch");
  }
  return fail(GRASPA_ERR_UNSUPPORTED, "widom: no launch shape fits the LDS (pair energies)");
}

// First-bead screen: the cell table in LDS when it fits, else read from L2.
template <bool CUBIC>
static int launch_first_bead(graspa_box* b, const WidomArgs& A, size_t base_lds, size_t tab_bytes)
{
  const bool tab_lds = A.use_cells && base_lds + tab_bytes + 20 * 1024 <= (size_t)b->max_lds;
  auto kern = tab_lds ? widom_first_bead_kernel<CUBIC, true> : widom_first_bead_kernel<CUBIC, false>;
  const size_t lds = tab_lds ? base_lds + tab_bytes : base_lds;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(kern, dim3((unsigned)A.groups), dim3(kFbWaves * 64), lds, b->stream, A);
  return hip_check(hipGetLastError(), "widom_first_bead_kernel launch");
}

// Scratch of one batch: the pipeline's masks, per-insertion state, live
// lists, work lists and their energies, sub-list lengths, work counts, the
// per-insertion records and the reduction partials.
struct WidomBufs {
  uint32_t* mask1;
  uint32_t* mask2;
  InsState* state;
  int* live;
  unsigned long long* stat_part;
  double2* e1;
  TrialPt* list1;
  double2* e2;
  TrialPt* list2;
  unsigned* counts;
  double* rec_w;
  double* rec_e;
  int* rec_f;
  double* part;
};

// Screening grid of a batch: up to kMaxScreenGroups workgroups, each owning
// a contiguous range of `chunk` insertions (four wave steps of kFbWaves x 64 /
// K insertions each); the work lists and the live list have a sub-list of
// chunk insertions per workgroup.  The first-bead, chain and finish kernels
// hold one workgroup per CU, so the group count is rounded down to a
// multiple of the CU count: a few groups beyond it would take a second round
// of the whole kernel.
static void widom_grid(const graspa_box* b, int K, int n_insert, int* groups, int* chunk)
{
  const long long per_group = 4LL * kFbWaves * (64 / K);
  long long raw = std::max<long long>(1, ((long long)n_insert + per_group - 1) / per_group);
  const long long ncu = std::max(1, b->n_cu);
  if (raw > ncu) raw = raw / ncu * ncu;
  const long long cap = kMaxScreenGroups >= ncu ? kMaxScreenGroups / ncu * ncu : kMaxScreenGroups;
  *groups = (int)std::min<long long>(raw, cap);
  *chunk = (n_insert + *groups - 1) / *groups;
}

// The box's own buffers, grown to the batch.
static int box_bufs(graspa_box* b, int n_insert, int K, int OC, int n_slots, WidomBufs* w)
{
  if (ensure_widom_records(b, n_insert)) return GRASPA_ERR_HIP;
  if (ensure_pipeline(b, n_slots, K, OC)) return GRASPA_ERR_HIP;
  w->mask1 = static_cast<uint32_t*>(b->d_wmask1);
  w->mask2 = static_cast<uint32_t*>(b->d_wmask2);
  w->state = static_cast<InsState*>(b->d_wstate);
  w->live = static_cast<int*>(b->d_wlive);
  w->stat_part = static_cast<unsigned long long*>(b->d_wstat);
  w->e1 = static_cast<double2*>(b->d_we1);
  w->list1 = static_cast<TrialPt*>(b->d_wlist1);
  w->e2 = static_cast<double2*>(b->d_we2);
  w->list2 = static_cast<TrialPt*>(b->d_wlist2);
  w->counts = static_cast<unsigned*>(b->d_wcounts);
  w->rec_w = b->d_wrec_w;
  w->rec_e = b->d_wrec_e;
  w->rec_f = b->d_wrec_f;
  w->part = b->d_wpart;
  return GRASPA_OK;
}

// Bytes of scratch a batch needs when every buffer is carved from one block
// (256-byte aligned pieces), and the carve-up itself when base is not null.
static size_t carve_bufs(char* base, int n_insert, int K, int OC, int n_slots, WidomBufs* w)
{
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  const size_t n = (size_t)n_slots;
  const size_t nblk = ((size_t)n_insert + kRedBlock * kRedPer - 1) / (kRedBlock * kRedPer);
  WidomBufs v;
  v.mask1 = reinterpret_cast<uint32_t*>(take(n * sizeof(uint32_t)));
  v.mask2 = reinterpret_cast<uint32_t*>(take(n * sizeof(uint32_t)));
  v.state = reinterpret_cast<InsState*>(take(n * sizeof(InsState)));
  v.live = reinterpret_cast<int*>(take(n * sizeof(int)));
  v.stat_part = reinterpret_cast<unsigned long long*>(
      take((2 * kMaxScreenGroups + 2 * kMaxEvalGroups) * sizeof(unsigned long long)));
  v.e1 = reinterpret_cast<double2*>(take(n * K * sizeof(double2)));
  v.list1 = reinterpret_cast<TrialPt*>(take(n * K * sizeof(TrialPt)));
  v.e2 = reinterpret_cast<double2*>(take(n * std::max(OC, 1) * sizeof(double2)));
  v.list2 = reinterpret_cast<TrialPt*>(take(n * std::max(OC, 1) * sizeof(TrialPt)));
  v.counts = reinterpret_cast<unsigned*>(take(3 * kMaxScreenGroups * sizeof(unsigned)));
  v.rec_w = reinterpret_cast<double*>(take((size_t)n_insert * sizeof(double)));
  v.rec_e = reinterpret_cast<double*>(take((size_t)n_insert * sizeof(double)));
  v.rec_f = reinterpret_cast<int*>(take((size_t)n_insert * sizeof(int)));
  v.part = reinterpret_cast<double*>(take(6 * std::max<size_t>(nblk, 1) * sizeof(double)));
  if (w) *w = v;
  return off;
}

// The no-allocation form's scratch: one block per device, its address taken
// once per device with hipGetSymbolAddress (host code never uses the
// symbol's name as an address).  Shared by every call on the device.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char widom_scratch[kWidomScratchBytes];
static std::mutex scratch_mu;
static char* scratch_ptr[kMaxDevices];   // per device, null until looked up

static int device_scratch(int device, char** out)
{
  if (device < 0 || device >= kMaxDevices) return fail(GRASPA_ERR_ARG, "widom scratch: device index out of range");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[device]) {
    void* p = nullptr;
    if (hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(widom_scratch)), "widom scratch address")) return GRASPA_ERR_HIP;
    scratch_ptr[device] = static_cast<char*>(p);
  }
  *out = scratch_ptr[device];
  return GRASPA_OK;
}

int launch_widom(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                 uint64_t seed, uint32_t stream, int64_t first_index, int n_insert, graspa_widom_result* d_res,
                 graspa_widom_summary* summary, const WidomLaunch& how)
{
  const bool no_alloc = how.no_alloc || how.args_only;
  if (comp < 1 || comp >= (int)b->comps.size()) return fail(GRASPA_ERR_ARG, "widom: bad component");
  const Component& c = b->comps[comp];
  if (K < 1 || K > kMaxTrials || O < 1 || O > kMaxTrials) return fail(GRASPA_ERR_ARG, "widom: 1 <= K,O <= 32");
  if (c.molsize < 1 || c.molsize > kMaxMol) return fail(GRASPA_ERR_UNSUPPORTED, "widom: molecule larger than 8 sites");
  if (c.d_chain)
    return fail(GRASPA_ERR_UNSUPPORTED, "widom: flexible chains take Widom moves in the batched driver (graspa_batch_*)");
  if (O * (c.molsize - 1) > kMaxPts) return fail(GRASPA_ERR_UNSUPPORTED, "widom: orientations x chain sites above 64");
  if (b->bc.ntypes > kMaxTypes) return fail(GRASPA_ERR_UNSUPPORTED, "widom: too many pseudo-atom types");
  if (n_insert <= 0) {
    if (summary) *summary = graspa_widom_summary{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    return GRASPA_OK;
  }
  const int chain = c.molsize - 1;
  int groups = 0, chunk = 0;
  widom_grid(b, K, n_insert, &groups, &chunk);
  WidomBufs W{};
  if (how.args_only) {
    const size_t ns = (size_t)groups * chunk, OC = (size_t)std::max(O * chain, 1);
    if (ns > b->wp_n || ns * K > b->wp_e1 || ns * OC > b->wp_e2 || (size_t)n_insert > b->wrec_cap)
      return fail(GRASPA_ERR_STATE, "widom: the box's buffers are smaller than the batch");
    W.mask1 = static_cast<uint32_t*>(b->d_wmask1);
    W.mask2 = static_cast<uint32_t*>(b->d_wmask2);
    W.state = static_cast<InsState*>(b->d_wstate);
    W.live = static_cast<int*>(b->d_wlive);
    W.stat_part = static_cast<unsigned long long*>(b->d_wstat);
    W.e1 = static_cast<double2*>(b->d_we1);
    W.list1 = static_cast<TrialPt*>(b->d_wlist1);
    W.e2 = static_cast<double2*>(b->d_we2);
    W.list2 = static_cast<TrialPt*>(b->d_wlist2);
    W.counts = static_cast<unsigned*>(b->d_wcounts);
    W.rec_w = b->d_wrec_w;
    W.rec_e = b->d_wrec_e;
    W.rec_f = b->d_wrec_f;
    W.part = b->d_wpart;
  } else if (how.no_alloc) {
    const size_t need = carve_bufs(nullptr, n_insert, K, O * chain, groups * chunk, nullptr);
    if (need > (size_t)kWidomScratchBytes)
      return fail(GRASPA_ERR_ARG, "widom (no allocation): the batch needs " + std::to_string(need) +
                                      " bytes of scratch, above the fixed " + std::to_string(kWidomScratchBytes) +
                                      "; use fewer insertions per call");
    char* base = nullptr;
    if (int rc = device_scratch(b->device, &base)) return rc;
    carve_bufs(base, n_insert, K, O * chain, groups * chunk, &W);
  } else if (int rc = box_bufs(b, n_insert, K, O * chain, groups * chunk, &W)) {
    return rc;
  }
  WidomArgs A;
  std::memset(&A, 0, sizeof(A));   // padding too: the argument bytes key a captured graph
  A.bc = b->bc;
  A.fw = b->d_fws; A.fw_type = b->d_fws_type; A.fw_fixT = b->d_fw_fixT; A.n_fw = b->n_fw;
  A.use_cells = b->use_cells;
  // Orientation sites: the cell screen costs about three times what it saves
  // in the pair kernel at 1,536 framework atoms (35 us of screening against
  // 12 us of pair energies per 100,000 insertions); the pair kernel's cost
  // grows with the framework and the screen's does not.
  A.screen_chain = b->use_cells && b->n_fw >= kChainScreenMinAtoms;
  for (int d = 0; d < 3; ++d) A.cell_n[d] = b->cell_n[d];
  A.cell_cap = b->cell_cap;
  A.ctab = b->d_ctab;
  A.g = b->d_g; A.g_type = b->d_g_type; A.g_molid = b->d_g_molid; A.g_fix = b->d_g_fix; A.guest_cap = b->guest_cap;
  guest_view(b, &A.gv);
  A.ff64 = b->d_ff64;
  A.comp = comp;
  A.new_molid = c.n_mol;
  A.M = c.molsize;
  A.chain = chain;
  for (int a = 0; a < c.molsize; ++a) {
    A.site_type[a] = c.type[a];
    A.site_q[a] = c.q[a];
    if (a >= 1)
      for (int d = 0; d < 3; ++d) A.tvec[a - 1][d] = c.pos[3 * a + d] - c.pos[d];
  }
  double qabs = 0.0;
  for (int a = 0; a < c.molsize; ++a) qabs += std::fabs(c.q[a]);
  A.charged = (b->bc.use_charges && qabs > 1e-6) ? 1 : 0;
  A.iso = (b->bc.cubic && b->bc.cell[0] == b->bc.cell[4] && b->bc.cell[0] == b->bc.cell[8]) ? 1 : 0;
  {
    // the fast pair loop's general form (see lj_general)
    bool general = std::any_of(b->shift.begin(), b->shift.end(), [](double x) { return x != 0.0; });
    const int T = b->bc.ntypes;
    for (int a = 0; a < c.molsize && !general; ++a)
      for (int t = 0; t < T && !general; ++t) {
        const int ij = t * T + c.type[a];
        const double s6 = std::pow(b->sig[ij] * b->sig[ij] / 0.01, 3.0);
        const double e01 = 4.0 * b->eps[ij] * (s6 * s6 - s6) - b->shift[ij];
        if (!(e01 > 2.0 * b->bc.overlap)) general = true;
      }
    A.lj_general = general ? 1 : 0;
  }
  A.K = K; A.O = O; A.beta = beta; A.bias = bias; A.excl = excl; A.tail_de = tail_de;
  A.seed = seed; A.stream = stream; A.first_index = how.first_dev ? 0 : first_index; A.n_insert = n_insert;
  A.first_dev = how.first_dev;
  A.n_kact = b->n_active;
  A.kact = b->d_kact;
  A.ktemp_act = b->d_ktemp_act;
  A.sk_ads = b->d_sk_ads; A.sk_fw = b->d_sk_fw;
  A.kstride = 1 + std::max(b->bc.kmax[0], std::max(b->bc.kmax[1], b->bc.kmax[2]));
  A.coul = b->d_coul;
  A.coul_n = b->coul_n;
  A.mask1 = W.mask1;
  A.mask2 = W.mask2;
  A.e1 = W.e1;
  A.e2 = W.e2;
  A.list1 = W.list1;
  A.list2 = W.list2;
  A.state = W.state;
  A.counts1 = W.counts;
  A.counts2 = A.counts1 + kMaxScreenGroups;
  A.live_counts = A.counts2 + kMaxScreenGroups;
  A.live = W.live;
  // list 1 and its counts are read by the first pair-energy launch only; the
  // chain kernel then writes the placed lists there for the finish kernel
  static_assert(sizeof(TrialPt) >= sizeof(int), "the placed lists fit in list 1");
  A.placed = reinterpret_cast<int*>(W.list1);
  A.placed_counts = A.counts1;
  A.stat_part = W.stat_part;
  for (int a = 0; a < c.molsize; ++a) {
    if (no_alloc) {
      // maps built by an earlier call are used; none is built here (a trial
      // a map would decide goes through the cell screen instead, with the
      // same outcome)
      const int t = c.type[a];
      A.dmap[a] = (t >= 0 && t < (int)b->d_dmap.size()) ? b->d_dmap[t] : nullptr;
    } else if (int rc = ensure_dead_map(b, c.type[a], &A.dmap[a])) {
      return rc;
    }
  }
  for (int d = 0; d < 3; ++d) A.dnv[d] = b->dmap_nv[d];
  {
    const int t0 = c.type[0];
    A.fmap0 = (A.dmap[0] && t0 >= 0 && t0 < (int)b->d_fmap.size()) ? b->d_fmap[t0] : nullptr;
  }
  A.groups = groups;
  A.chunk = chunk;
  A.res = d_res;
  A.rec_w = W.rec_w; A.rec_e = W.rec_e; A.rec_f = W.rec_f;
  A.stat = reinterpret_cast<unsigned long long*>(b->d_wacc + 8);
  const int T = A.bc.ntypes, M = A.M;
  // screening kernels: thresholds | cell table (optional) | per-wave areas
  A.lds_thr = 0;
  const size_t scr_base = align16((size_t)M * T * sizeof(float));
  A.lds_tab = (int)scr_base;
  const size_t tab_bytes =
      A.use_cells ? (size_t)b->cell_n[0] * b->cell_n[1] * b->cell_n[2] * b->cell_cap * sizeof(uint4) : 0;
  // pair-energy kernel: Coulomb table (fast form) | FF rows | framework (optional)
  A.lds_coul = 0;
  A.lds_ffl = (int)eval_lds_coul_bytes(b);
  A.lds_fw = (int)eval_lds_base(b, M);
  const bool cubic = A.bc.cubic != 0;
  if (how.key) {
    const unsigned char* a = reinterpret_cast<const unsigned char*>(&A);
    const int shape[4] = {b->ev_waves, b->ev_per_cu, b->ev_fw_lds, groups};
    const unsigned char* sh = reinterpret_cast<const unsigned char*>(shape);
    how.key->assign(a, a + sizeof(A));
    how.key->insert(how.key->end(), sh, sh + sizeof(shape));
  }
  if (how.args_only) return GRASPA_OK;

  int rc = cubic ? launch_first_bead<true>(b, A, scr_base, tab_bytes) : launch_first_bead<false>(b, A, scr_base, tab_bytes);
  if (rc) return rc;
  int egrid[2] = {0, 0};
  auto eval = [&](const TrialPt* list, const unsigned* counts, int per_ins, double2* energy, uint32_t* mask, int e) {
    unsigned long long* slot = A.stat_part + 2 * kMaxScreenGroups + (size_t)e * kMaxEvalGroups;
    return launch_eval(b, A, list, counts, A.groups, A.chunk, per_ins, energy, mask, slot, &egrid[e],
                       !no_alloc && !how.first_dev);
  };
  rc = eval(A.list1, A.counts1, K, A.e1, A.mask1, 0);
  if (rc) return rc;
  rc = launch_ins(b, cubic ? widom_chain_kernel<true> : widom_chain_kernel<false>, A, scr_base, "widom_chain_kernel",
                  kInsWaves, A.groups);
  if (rc) return rc;
  if (chain > 0) {
    rc = eval(A.list2, A.counts2, O * chain, A.e2, A.mask2, 1);
    if (rc) return rc;
  }
  const size_t fin_lds = (size_t)kFinWaves * ((size_t)M * 3 * A.kstride * sizeof(double2) + kMaxMol * 3 * sizeof(double));
  {
    // several workgroups per placed list (as many as stay resident), so its
    // insertions spread over more waves than one workgroup's
    const int occ = occupancy(b, widom_finish_kernel, kFinWaves * 64, fin_lds);
    const long long slots = (long long)std::max(1, occ) * std::max(1, b->n_cu);
    const int per = (int)std::max(1LL, std::min<long long>(kFinMaxPer, slots / A.groups));
    rc = launch_ins(b, widom_finish_kernel, A, fin_lds, "widom_finish_kernel", kFinWaves, A.groups * per);
  }
  if (rc) return rc;
  const int nblk = (n_insert + kRedBlock * kRedPer - 1) / (kRedBlock * kRedPer);
  hipLaunchKernelGGL(widom_reduce_kernel, dim3(nblk), dim3(kRedBlock), 0, b->stream, A.rec_w, A.rec_e, A.rec_f,
                     n_insert, W.part);
  if (hip_check(hipGetLastError(), "widom_reduce launch")) return GRASPA_ERR_HIP;
  hipLaunchKernelGGL(widom_final_kernel, dim3(1), dim3(kRedBlock), 0, b->stream, W.part, nblk, b->d_wacc,
                     b->d_wacc + kWaccTotals, A.stat_part, A.groups, egrid[0], egrid[1], A.stat,
                     reinterpret_cast<unsigned long long*>(b->d_wacc + kWaccTotals + 8));
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

// Launch-shape autotuner of the pair-energy kernel.  Every shape that fits
// (4, 8, 12 or 16 waves per workgroup; framework staged in LDS or read from
// L2; 1 up to the occupancy, at most kMaxTunePerCu, workgroups per CU) runs
// the given Widom batch `reps` times after one untimed run, timed with HIP
// events on the box's stream, and the box keeps the fastest.  A shape only
// decides which wave evaluates which trial site (one wave per site, lanes
// over the same atoms in the same order), so every shape gives the same
// results bit for bit.
int widom_autotune(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                   uint64_t seed, int n_insert, int reps, graspa_eval_shape* tried, double* ms, int cap, int* n_tried)
{
  if (n_insert < 1 || reps < 1 || cap < 0) return fail(GRASPA_ERR_ARG, "widom autotune: n_insert >= 1, reps >= 1");
  if (comp < 1 || comp >= (int)b->comps.size()) return fail(GRASPA_ERR_ARG, "widom: bad component");
  const size_t base = eval_lds_base(b, b->comps[comp].molsize);
  std::vector<graspa_eval_shape> cand;
  for (const int w : {4, 8, 12, 16})
    for (int f = 1; f >= 0; --f) {
      const int occ = occupancy(b, eval_kernel_for(b, w, f != 0), w * 64, eval_lds(b, base, f != 0));
      for (int p = 1; p <= std::min(occ, kMaxTunePerCu); ++p) cand.push_back(graspa_eval_shape{w, p, f});
    }
  if (cand.empty()) return fail(GRASPA_ERR_UNSUPPORTED, "widom: no launch shape fits the LDS (pair energies)");
  const graspa_eval_shape keep{b->ev_waves, b->ev_per_cu, b->ev_fw_lds};
  // the timed batches add into the running totals and overwrite the last
  // batch's summary (d_wacc): both are saved here and restored at the end, so
  // a tuning run between queued production batches leaves them as they were
  double saved_acc[kWaccSize];
  int rc = hip_check(hipMemcpyAsync(saved_acc, b->d_wacc, sizeof(saved_acc), hipMemcpyDeviceToHost, b->stream),
                     "widom autotune save totals");
  rc = rc ? rc : hip_check(hipStreamSynchronize(b->stream), "widom autotune save totals");
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = hip_check(hipEventCreate(&e0), "widom autotune event");
  rc = rc ? rc : hip_check(hipEventCreate(&e1), "widom autotune event");
  double best = 0.0;
  size_t ibest = 0;
  for (size_t i = 0; i < cand.size() && !rc; ++i) {
    b->ev_waves = cand[i].waves;
    b->ev_per_cu = cand[i].per_cu;
    b->ev_fw_lds = cand[i].framework_lds;
    rc = launch_widom(b, comp, K, O, beta, bias, excl, tail_de, seed, 0u, 0, n_insert, nullptr, nullptr, WidomLaunch());
    rc = rc ? rc : hip_check(hipEventRecord(e0, b->stream), "widom autotune event");
    for (int r = 0; r < reps && !rc; ++r)
      rc = launch_widom(b, comp, K, O, beta, bias, excl, tail_de, seed, 0u, 0, n_insert, nullptr, nullptr, WidomLaunch());
    rc = rc ? rc : hip_check(hipEventRecord(e1, b->stream), "widom autotune event");
    rc = rc ? rc : hip_check(hipEventSynchronize(e1), "widom autotune");
    float t = 0.0f;
    rc = rc ? rc : hip_check(hipEventElapsedTime(&t, e0, e1), "widom autotune timing");
    if (rc) break;
    const double per = (double)t / reps;
    if ((int)i < cap) {
      if (tried) tried[i] = cand[i];
      if (ms) ms[i] = per;
    }
    if (i == 0 || per < best) {
      best = per;
      ibest = i;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  int rc_restore = hip_check(hipMemcpyAsync(b->d_wacc, saved_acc, sizeof(saved_acc), hipMemcpyHostToDevice, b->stream),
                             "widom autotune restore totals");
  rc_restore = rc_restore ? rc_restore : hip_check(hipStreamSynchronize(b->stream), "widom autotune restore totals");
  rc = rc ? rc : rc_restore;
  const graspa_eval_shape& pick = rc ? keep : cand[ibest];
  b->ev_waves = pick.waves;
  b->ev_per_cu = pick.per_cu;
  b->ev_fw_lds = pick.framework_lds;
  if (rc) return rc;
  if (n_tried) *n_tried = (int)cand.size();
  return GRASPA_OK;
}

// Advances a captured batch's first insertion index past the batch (the
// last node of the graph).
__global__ void widom_advance_kernel(int64_t* first, int n)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) *first += n;
}

void widom_graph_free(graspa_box* b)
{
  if (b->wg_exec) (void)hipGraphExecDestroy(b->wg_exec);
  if (b->wg_graph) (void)hipGraphDestroy(b->wg_graph);
  if (b->d_wg_first) (void)box_free(b, b->d_wg_first);
  b->wg_exec = nullptr;
  b->wg_graph = nullptr;
  b->d_wg_first = nullptr;
  b->wg_key.clear();
  b->wg_n = 0;
}

// Captures one queued batch (launch_widom with its first index in device
// memory, then the advance kernel) from the box's stream into a graph.  The
// box's buffers and voxel maps are grown and built first by an uncaptured
// argument pass, so nothing is allocated while the stream is captured.
int widom_graph_create(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                       uint64_t seed, uint32_t stream, int64_t first_index, int n_insert)
{
  if (n_insert < 1) return fail(GRASPA_ERR_ARG, "widom graph: n_insert >= 1");
  if (hip_check(hipStreamSynchronize(b->stream), "widom graph")) return GRASPA_ERR_HIP;
  widom_graph_free(b);
  if (comp < 1 || comp >= (int)b->comps.size()) return fail(GRASPA_ERR_ARG, "widom: bad component");
  const Component& c = b->comps[comp];
  int groups = 0, chunk = 0;
  if (K < 1 || K > kMaxTrials) return fail(GRASPA_ERR_ARG, "widom: 1 <= K,O <= 32");
  widom_grid(b, K, n_insert, &groups, &chunk);
  WidomBufs W{};
  if (int rc = box_bufs(b, n_insert, K, O * std::max(c.molsize - 1, 0), groups * chunk, &W)) return rc;
  for (int a = 0; a < c.molsize; ++a) {
    const uint32_t* m = nullptr;
    if (int rc = ensure_dead_map(b, c.type[a], &m)) return rc;
  }
  if (hip_check(box_malloc(b, &b->d_wg_first, sizeof(int64_t)), "hipMalloc widom graph")) return GRASPA_ERR_HIP;
  if (hip_check(hipMemcpy(b->d_wg_first, &first_index, sizeof(int64_t), hipMemcpyHostToDevice), "widom graph")) {
    widom_graph_free(b);
    return GRASPA_ERR_HIP;
  }
  WidomLaunch how;
  how.first_dev = b->d_wg_first;
  how.key = &b->wg_key;
  if (hip_check(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "widom graph capture")) {
    widom_graph_free(b);
    return GRASPA_ERR_HIP;
  }
  int rc = launch_widom(b, comp, K, O, beta, bias, excl, tail_de, seed, stream, first_index, n_insert, nullptr, nullptr,
                        how);
  if (!rc) {
    hipLaunchKernelGGL(widom_advance_kernel, dim3(1), dim3(64), 0, b->stream, b->d_wg_first, n_insert);
    rc = hip_check(hipGetLastError(), "widom_advance launch");
  }
  hipGraph_t g = nullptr;
  const int rc_end = hip_check(hipStreamEndCapture(b->stream, &g), "widom graph capture");
  b->wg_graph = g;
  rc = rc ? rc : rc_end;
  if (!rc) rc = hip_check(hipGraphInstantiate(&b->wg_exec, b->wg_graph, nullptr, nullptr, 0), "widom graph instantiate");
  if (rc) {
    widom_graph_free(b);
    return rc;
  }
  b->wg_comp = comp; b->wg_K = K; b->wg_O = O; b->wg_bias = bias; b->wg_n = n_insert;
  b->wg_beta = beta; b->wg_excl = excl; b->wg_tail = tail_de; b->wg_seed = seed; b->wg_stream = stream;
  return GRASPA_OK;
}

// Queues n replays of the captured batch, after checking that the box's
// kernel arguments for that batch are still the captured ones (molecules,
// structure-factor buffer, framework, buffers, launch shape).
int widom_graph_launch(graspa_box* b, int n_replays)
{
  if (!b->wg_exec) return fail(GRASPA_ERR_STATE, "widom graph: none captured");
  if (n_replays < 0) return fail(GRASPA_ERR_ARG, "widom graph: n_replays >= 0");
  std::vector<unsigned char> key;
  WidomLaunch how;
  how.first_dev = b->d_wg_first;
  how.args_only = true;
  how.key = &key;
  int rc = launch_widom(b, b->wg_comp, b->wg_K, b->wg_O, b->wg_beta, b->wg_bias, b->wg_excl, b->wg_tail, b->wg_seed,
                        b->wg_stream, 0, b->wg_n, nullptr, nullptr, how);
  if (rc || key != b->wg_key)
    return fail(GRASPA_ERR_STATE, "widom graph: the box changed since the capture; capture again");
  for (int r = 0; r < n_replays; ++r)
    if (hip_check(hipGraphLaunch(b->wg_exec, b->stream), "widom graph launch")) return GRASPA_ERR_HIP;
  return GRASPA_OK;
}

}  // namespace graspa
