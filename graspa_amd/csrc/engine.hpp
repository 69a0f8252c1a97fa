// engine.hpp -- internal state of a box resident on one MI355X and the
// launchers of the engine's kernels.  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/graspa_engine.h"

namespace graspa {

// Sites passed to the per-move kernels by value in the kernel argument block:
// one CBMC trial batch or one moved molecule (old + new) fits.
constexpr int kMaxSites = 64;
constexpr int kMaxTypes = 32;
static_assert(kMaxTypes == GRASPA_MAX_PSEUDO, "pseudo-atom counts cover every type a box may have");
constexpr int kWaccTotals = 16;    // offset of the running Widom totals in graspa_box::d_wacc
constexpr int kWaccSize = 32;      // doubles of graspa_box::d_wacc
constexpr int kMaxComps = 8;
// flexible chain model of a species: per bead i < kChainBeads, {bond,
// theta0, k_bend, t0, t1, t2, t3} (graspa_chain, packed)
constexpr int kChainBeads = 4;
constexpr int kChainPer = 7;
constexpr int kChainDoubles = kChainBeads * kChainPer;
constexpr int kBlock = 256;          // 4 wave64 per workgroup

struct SiteBatch {
  int n;
  int _pad;
  double x[kMaxSites], y[kMaxSites], z[kMaxSites], q[kMaxSites];
  int type[kMaxSites];
};

// Geometry and force field as the kernels see them.
struct BoxConst {
  double cell[9], inv[9];
  int cubic;
  int ntypes;
  double cut_vdw2, cut_coul2, overlap;
  double alpha, prefactor;
  int use_charges;
  int kmax[3];
  int nvec;
  // mixed-precision helpers: fixed-point fraction -> Cartesian, per column
  float h32[9];                      // cell * 2^-32
  double h64[9];                     // cell * 2^-32 (FP64)
  float cut_vdw2f, cut_coul2f, overlapf, alphaf, preff;
};

// Guest segment of one component in the combined guest store.
struct CompSeg {
  int base;      // first atom
  int n_atoms;   // n_mol * molsize
};

struct GuestView {
  int n_comp;
  int total;     // atoms in use over all components (for bounds)
  CompSeg seg[kMaxComps];   // index 1..n_comp-1 used, 0 unused (framework)
};

struct Component {
  int molsize = 0, capacity = 0, n_mol = 0, base = 0;
  std::vector<double> pos;      // host mirror, capacity*molsize*3
  std::vector<double> q;
  std::vector<int> type;
  std::vector<int64_t> molid;
  // flexible chain model on the device (kChainDoubles doubles, move_math.hpp
  // layout); null for a rigid molecule
  double* d_chain = nullptr;
};

struct TimingState {
  std::string family;
  bool on = false;
  std::vector<hipEvent_t> ev;        // pairs
  int64_t launches = 0;
};

}  // namespace graspa

struct graspa_box {
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = GRASPA_FP64;
  graspa::BoxConst bc{};
  double volume = 0.0;
  double recip_cut = 0.0;
  int lammps_ewald = 0;
  std::vector<double> eps, sig, shift;   // host copies [T*T]

  // force field tables: {4 eps, sigma^2, shift, 0} per type pair
  double* d_ff64 = nullptr;
  float* d_ff32 = nullptr;

  // framework (component 0)
  int n_fw = 0;
  double* d_fw = nullptr;       // SoA [5][n_fw]: x, y, z, q, (type as double unused)
  int* d_fw_type = nullptr;
  uint4* d_fw_fix = nullptr;    // fixed-point fraction x,y,z + float q bits
  uint4* d_fw_fixT = nullptr;   // fixed-point fraction x,y,z + type
  // Widom copy of the framework in Morton order of its screen cells, and the
  // screen's cell table (cap slots per cell, empty slots have type ~0u)
  double* d_fws = nullptr;      // SoA [4][n_fw] x, y, z, q
  int* d_fws_type = nullptr;
  uint4* d_ctab = nullptr;      // [n_cells * cap] fixed-point x,y,z + type
  int cell_n[3] = {0, 0, 0};
  int cell_cap = 0;
  int use_cells = 0;
  // maps of the framework's certain-overlap volume per probe-site type
  // (dead_map.cpp), built on first use
  std::vector<uint32_t*> d_dmap;
  std::vector<uint32_t*> d_fmap;         // certainly free voxels, same layout (built with d_dmap)
  int dmap_nv[3] = {0, 0, 0};
  // real-space Coulomb table of the fast form (coulomb_table.cpp)
  double* d_coul = nullptr;
  int coul_n = 0;               // doubles
  std::vector<double> fw_pos, fw_q;
  std::vector<int> fw_type;

  // guests: one store, a segment per component
  std::vector<graspa::Component> comps;  // [0] = framework placeholder
  int guest_cap = 0;
  double* d_g = nullptr;        // SoA [4][guest_cap]: x, y, z, q
  int* d_g_type = nullptr;
  int* d_g_molid = nullptr;
  int* d_g_comp = nullptr;
  uint4* d_g_fix = nullptr;

  // Ewald
  int nvec = 0, n_active = 0;
  double* d_ktemp = nullptr;    // [nvec] prefactor * exp(-k^2/4a^2)/k^2 (x2 for kx>0), 0 if inactive
  double2* d_sk_fw = nullptr;
  double2* d_sk_ads = nullptr;
  double2* d_sk_stage = nullptr;
  int4* d_kact = nullptr;       // active k-vectors: kx, ky, kz, index
  double* d_ktemp_act = nullptr;  // ktemp of the active k-vectors
  std::vector<double> ktemp;
  bool sk_staged = false;

  // scratch for per-move results
  double* d_part = nullptr;     // block partials
  size_t part_cap = 0;
  double* h_part = nullptr;     // pinned mirror
  unsigned* d_ticket = nullptr;

  // Widom scratch
  double* d_wacc = nullptr;     // [0,6) batch sums, [8,10) work counters, [12] insertion counter;
                                // from kWaccTotals the same sums and counters over all batches since the last reset
  double* d_wrec_w = nullptr;   // per-insertion W, dE, flags
  double* d_wrec_e = nullptr;
  int* d_wrec_f = nullptr;
  double* d_wpart = nullptr;    // reduction partials
  size_t wrec_cap = 0;
  // Widom pipeline buffers (trial masks, energies, work lists, per-insertion state)
  void* d_wmask1 = nullptr;
  void* d_wmask2 = nullptr;
  void* d_we1 = nullptr;
  void* d_we2 = nullptr;
  void* d_wlist1 = nullptr;
  void* d_wlist2 = nullptr;
  void* d_wstate = nullptr;
  void* d_wcounts = nullptr;    // sub-list lengths of work lists 1 and 2 and of the live lists
  void* d_wlive = nullptr;      // insertions with a first-bead trial left after the screen
  void* d_wstat = nullptr;      // per-workgroup work counts of the pipeline
  size_t wp_n = 0, wp_e1 = 0, wp_e2 = 0;
  graspa_widom_result* d_wres = nullptr;
  size_t wres_cap = 0;
  int n_cu = 256;
  int max_lds = 65536;          // dynamic LDS per workgroup
  // launch shape of the Widom pair-energy kernel (graspa_widom_set_shape,
  // graspa_widom_autotune): waves per workgroup, workgroups per CU (0: as
  // many as are resident), framework staged in LDS when it fits
  // a captured Widom batch (graspa_widom_graph_create): the instantiated
  // graph, the device-side first insertion index its replays advance, and
  // the bytes of the captured arguments (a replay is refused when the box's
  // arguments no longer match them)
  hipGraph_t wg_graph = nullptr;
  hipGraphExec_t wg_exec = nullptr;
  int64_t* d_wg_first = nullptr;
  std::vector<unsigned char> wg_key;
  int wg_comp = 0, wg_K = 0, wg_O = 0, wg_bias = 0, wg_n = 0;
  double wg_beta = 0.0, wg_excl = 0.0, wg_tail = 0.0;
  uint64_t wg_seed = 0;
  uint32_t wg_stream = 0;
  int ev_waves = 12;            // two 12-wave workgroups per CU with the framework in LDS
  int ev_per_cu = 0;
  int ev_fw_lds = 1;

  graspa::TimingState timing;
  // device allocations and their total (box_malloc / box_free)
  std::unordered_map<void*, size_t> dev_alloc;
  size_t dev_bytes = 0;
};

// Device allocations of a box, recorded so that graspa_box_footprint can
// report what the box holds.
template <typename T>
inline hipError_t box_malloc(graspa_box* b, T** p, size_t bytes)
{
  const hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  if (e == hipSuccess && *p) {
    b->dev_alloc[static_cast<void*>(*p)] = bytes;
    b->dev_bytes += bytes;
  }
  return e;
}

inline void box_free(graspa_box* b, void* p)
{
  if (!p) return;
  auto it = b->dev_alloc.find(p);
  if (it != b->dev_alloc.end()) {
    b->dev_bytes -= it->second;
    b->dev_alloc.erase(it);
  }
  (void)hipFree(p);
}

namespace graspa {

// error reporting
int fail(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);

// geometry helpers shared by host code
void fixed_from_cart(const BoxConst& bc, const double* p, uint32_t out[3]);
void guest_view(const graspa_box* b, GuestView* gv);

// kernel launchers (engine_kernels.hip)
int launch_site_energies(graspa_box* b, const SiteBatch& sb, int self_comp, int self_molid, int excl_comp,
                         int excl_molid, double* d_out /* n_sites*5 */);
int launch_ewald_delta(graspa_box* b, const SiteBatch& sb, int n_old, double* d_out /* blocks*2 */, int* n_blocks);
int launch_ewald_build(graspa_box* b, double* d_out_energy /* blocks*3 */, int* n_blocks);
int launch_total_pairs(graspa_box* b, double* d_out /* blocks*4 */, int* n_blocks);

// writes atoms [first_atom, first_atom + n_atoms) of a component from the host
// mirror into the device guest store (engine_api.hip)
int upload_guest_atoms(graspa_box* b, int comp, int first_atom, int n_atoms);

// per-insertion record buffers of the Widom kernel (engine_api.hip)
int ensure_widom_records(graspa_box* b, int n_insert);

// Squared-distance threshold below which a pair of these force-field
// parameters certainly overlaps (energy above the criterion or r^2 < 0.01,
// inside the VDW cutoff); used by the Widom overlap screen and its cell grid.
__host__ __device__ inline double overlap_r2(double a4e, double s2, double shift, double overlap, double cut2)
{
  double thr = 0.01;
  if (a4e > 0.0 && s2 > 0.0) {
    const double c = (overlap + shift) / a4e;
    if (c > 0.0) {
      const double xp = 0.5 * (1.0 + sqrt(1.0 + 4.0 * c));
      const double r2 = s2 / cbrt(xp);
      thr = thr > r2 ? thr : r2;
    }
  }
  return thr < cut2 ? thr : cut2;
}

// screen cell grid of the framework (engine_api.hip)
int build_screen_cells(graspa_box* b);

// Certain-overlap voxel maps of the framework (dead_map.cpp): voxels of about
// kDeadVoxel A; bit (x + nv0 (y + nv1 z)) set when every point of the voxel
// overlaps a framework atom for a probe site of the map's type.
constexpr double kDeadVoxel = 0.15;
void dead_map_dims(const graspa_box* b, int nv[3]);
void build_dead_map(const graspa_box* b, int site_type, std::vector<uint32_t>& bits, int nv[3],
                    std::vector<uint32_t>* free_bits = nullptr);
int ensure_dead_map(graspa_box* b, int site_type, const uint32_t** out);
void free_dead_maps(graspa_box* b);

// Real-space Coulomb table g(r^2) = erfc(alpha r)/r of the fast pair form
// (coulomb_table.cpp): kCoulTabM intervals per octave of r^2 from
// 2^kCoulTabELo, a degree-kCoulTabDeg polynomial in r^2 - r_i^2 per interval.
// 16 intervals per octave with degree 7 keep the table at 15 KB for a 12 A
// cutoff (64 intervals with degree 5 took 46 KB at the same accuracy), so
// the pair kernels' LDS leaves room for a second workgroup per CU.
constexpr int kCoulTabELo = -7;
constexpr int kCoulTabMBits = 4;
constexpr int kCoulTabM = 1 << kCoulTabMBits;
constexpr int kCoulTabDeg = 7;
constexpr int kCoulTabStride = 8;    // doubles per interval (kCoulTabDeg + 1)
static_assert(kCoulTabStride == kCoulTabDeg + 1 && kCoulTabStride % 2 == 0, "table rows are read as double2");
// The pair kernels stage the table in LDS with rows kCoulLdsRow double2 (80 B)
// apart, the last double2 of a row unused.  Lanes of a wave read rows of
// different intervals at once: with rows 64 B apart a given coefficient of
// every interval falls in one of only two 16-byte bank groups of the 32 banks,
// with 80 B rows in one of eight.
constexpr int kCoulLdsRow = 5;
static_assert(kCoulLdsRow >= kCoulTabStride / 2, "an LDS row holds a table row");
// doubles of the LDS copy of a table of coul_n doubles
constexpr int coul_lds_doubles(int coul_n) { return coul_n / kCoulTabStride * 2 * kCoulLdsRow; }
int coulomb_table_octaves(double cut2);
void build_coulomb_table(double alpha, double cut2, std::vector<double>& tab);

// widom_kernel.hip
// Bytes of the no-allocation form's fixed scratch block (one per device).
constexpr size_t kWidomScratchBytes = (size_t)32 << 20;
// How a Widom batch is launched.  no_alloc: the pipeline's scratch comes
// from that block, and no buffer or voxel map is allocated (maps built by
// earlier calls are used).  first_dev: the batch's first insertion index is
// read from device memory (graph replays) instead of first_index.
// args_only: nothing is launched or allocated; key receives the bytes of the
// kernels' arguments and launch shape (the box's buffers as they are; an
// error if they are too small).
struct WidomLaunch {
  bool no_alloc = false;
  const int64_t* first_dev = nullptr;
  bool args_only = false;
  std::vector<unsigned char>* key = nullptr;
};
int launch_widom(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                 uint64_t seed, uint32_t stream, int64_t first_index, int n_insert, graspa_widom_result* d_res,
                 graspa_widom_summary* summary, const WidomLaunch& how);
int widom_graph_create(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                       uint64_t seed, uint32_t stream, int64_t first_index, int n_insert);
int widom_graph_launch(graspa_box* b, int n_replays);
void widom_graph_free(graspa_box* b);
int widom_autotune(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                   uint64_t seed, int n_insert, int reps, graspa_eval_shape* tried, double* ms, int cap, int* n_tried);

// timing
void timing_begin(graspa_box* b, const char* family);
void timing_end(graspa_box* b, const char* family);

}  // namespace graspa
