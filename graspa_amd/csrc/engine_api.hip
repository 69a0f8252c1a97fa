// engine_api.hip -- the C-ABI of the GCMC energy engine: box life cycle, state
// uploads, per-move energy entry points and measurement hooks.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "device_math.hpp"
#include "engine.hpp"

namespace graspa {

static thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
  g_last_error = msg;
  return code;
}

int hip_check(hipError_t e, const char* what)
{
  if (e == hipSuccess) return GRASPA_OK;
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return GRASPA_ERR_HIP;
}

void fixed_from_cart(const BoxConst& bc, const double* p, uint32_t out[3])
{
  double s[3];
  s[0] = bc.inv[0] * p[0] + bc.inv[3] * p[1] + bc.inv[6] * p[2];
  s[1] = bc.inv[1] * p[0] + bc.inv[4] * p[1] + bc.inv[7] * p[2];
  s[2] = bc.inv[2] * p[0] + bc.inv[5] * p[1] + bc.inv[8] * p[2];
  for (int d = 0; d < 3; ++d) {
    const double fr = s[d] - std::floor(s[d]);
    out[d] = (uint32_t)(int64_t)(fr * 4294967296.0);
  }
}

void guest_view(const graspa_box* b, GuestView* gv)
{
  std::memset(gv, 0, sizeof(*gv));
  gv->n_comp = (int)b->comps.size();
  int total = 0;
  for (int c = 1; c < gv->n_comp; ++c) {
    gv->seg[c].base = b->comps[c].base;
    gv->seg[c].n_atoms = b->comps[c].n_mol * b->comps[c].molsize;
    total += gv->seg[c].n_atoms;
  }
  gv->total = total;
}

void timing_begin(graspa_box* b, const char* family)
{
  if (!b->timing.on || b->timing.family != family) return;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, b->stream);
  b->timing.ev.push_back(e0);
  b->timing.ev.push_back(e1);
}

void timing_end(graspa_box* b, const char* family)
{
  if (!b->timing.on || b->timing.family != family || b->timing.ev.empty()) return;
  hipEventRecord(b->timing.ev.back(), b->stream);
  b->timing.launches++;
}

static void timing_clear(graspa_box* b)
{
  for (auto e : b->timing.ev) hipEventDestroy(e);
  b->timing.ev.clear();
  b->timing.launches = 0;
}

// Writes up to kMaxSites atoms of the guest store (commit of one molecule).
struct AtomPatch {
  int n;
  int idx[kMaxSites];
  double x[kMaxSites], y[kMaxSites], z[kMaxSites], q[kMaxSites];
  int type[kMaxSites], molid[kMaxSites], comp[kMaxSites];
  uint4 fix[kMaxSites];
};

__global__ void patch_atoms_kernel(const AtomPatch p, double* g, int cap, int* g_type, int* g_molid, int* g_comp,
                                   uint4* g_fix)
{
  const int i = threadIdx.x;
  if (i >= p.n) return;
  const int j = p.idx[i];
  g[j] = p.x[i];
  g[cap + j] = p.y[i];
  g[2 * cap + j] = p.z[i];
  g[3 * cap + j] = p.q[i];
  g_type[j] = p.type[i];
  g_molid[j] = p.molid[i];
  g_comp[j] = p.comp[i];
  g_fix[j] = p.fix[i];
}

}  // namespace graspa

using namespace graspa;

int graspa::upload_guest_atoms(graspa_box* b, int comp, int first_atom, int n_atoms)
{
  // one launch per kMaxSites atoms, from the host mirror
  const Component& c = b->comps[comp];
  for (int off = 0; off < n_atoms; off += kMaxSites) {
    AtomPatch p{};
    p.n = std::min(kMaxSites, n_atoms - off);
    for (int i = 0; i < p.n; ++i) {
      const int a = first_atom + off + i;
      p.idx[i] = c.base + a;
      p.x[i] = c.pos[3 * a];
      p.y[i] = c.pos[3 * a + 1];
      p.z[i] = c.pos[3 * a + 2];
      p.q[i] = c.q[a];
      p.type[i] = c.type[a];
      p.molid[i] = (int)c.molid[a];
      p.comp[i] = comp;
      uint32_t f[3];
      fixed_from_cart(b->bc, &c.pos[3 * a], f);
      float qf = (float)c.q[a];
      uint32_t qb;
      std::memcpy(&qb, &qf, 4);
      p.fix[i] = make_uint4(f[0], f[1], f[2], qb);
    }
    hipLaunchKernelGGL(patch_atoms_kernel, dim3(1), dim3(kMaxSites), 0, b->stream, p, b->d_g, b->guest_cap,
                       b->d_g_type, b->d_g_molid, b->d_g_comp, b->d_g_fix);
    if (hip_check(hipGetLastError(), "patch_atoms launch")) return GRASPA_ERR_HIP;
  }
  return GRASPA_OK;
}

static int realloc_guests(graspa_box* b)
{
  int cap = 0;
  for (size_t c = 1; c < b->comps.size(); ++c) {
    b->comps[c].base = cap;
    cap += b->comps[c].capacity * b->comps[c].molsize;
  }
  box_free(b, b->d_g); box_free(b, b->d_g_type); box_free(b, b->d_g_molid); box_free(b, b->d_g_comp); box_free(b, b->d_g_fix);
  b->d_g = nullptr; b->d_g_type = b->d_g_molid = b->d_g_comp = nullptr; b->d_g_fix = nullptr;
  b->guest_cap = cap;
  const size_t n = (size_t)std::max(cap, 1);
  if (hip_check(box_malloc(b, &b->d_g, 4 * n * sizeof(double)), "hipMalloc guests")) return GRASPA_ERR_HIP;
  if (hip_check(box_malloc(b, &b->d_g_type, n * sizeof(int)), "hipMalloc guests")) return GRASPA_ERR_HIP;
  if (hip_check(box_malloc(b, &b->d_g_molid, n * sizeof(int)), "hipMalloc guests")) return GRASPA_ERR_HIP;
  if (hip_check(box_malloc(b, &b->d_g_comp, n * sizeof(int)), "hipMalloc guests")) return GRASPA_ERR_HIP;
  if (hip_check(box_malloc(b, &b->d_g_fix, n * sizeof(uint4)), "hipMalloc guests")) return GRASPA_ERR_HIP;
  hipMemsetAsync(b->d_g, 0, 4 * n * sizeof(double), b->stream);
  hipMemsetAsync(b->d_g_type, 0, n * sizeof(int), b->stream);
  hipMemsetAsync(b->d_g_molid, 0xff, n * sizeof(int), b->stream);
  hipMemsetAsync(b->d_g_comp, 0, n * sizeof(int), b->stream);
  hipMemsetAsync(b->d_g_fix, 0, n * sizeof(uint4), b->stream);
  for (size_t c = 1; c < b->comps.size(); ++c) {
    const Component& cc = b->comps[c];
    // slot 0 holds the template while the component is empty (as in the reference)
    const int n_atoms = std::max(cc.n_mol, 1) * cc.molsize;
    if (int rc = upload_guest_atoms(b, (int)c, 0, n_atoms)) return rc;
  }
  return hip_check(hipStreamSynchronize(b->stream), "guest upload");
}

static int ensure_part(graspa_box* b, size_t n_doubles)
{
  if (n_doubles <= b->part_cap) return GRASPA_OK;
  box_free(b, b->d_part);
  hipHostFree(b->h_part);
  b->d_part = nullptr;
  b->h_part = nullptr;
  size_t cap = std::max<size_t>(n_doubles, 4096);
  if (hip_check(box_malloc(b, &b->d_part, cap * sizeof(double)), "hipMalloc partials")) return GRASPA_ERR_HIP;
  if (hip_check(hipHostMalloc((void**)&b->h_part, cap * sizeof(double), hipHostMallocDefault), "hipHostMalloc"))
    return GRASPA_ERR_HIP;
  b->part_cap = cap;
  return GRASPA_OK;
}

int graspa::ensure_widom_records(graspa_box* b, int n_insert)
{
  if ((size_t)n_insert <= b->wrec_cap) return GRASPA_OK;
  hipStreamSynchronize(b->stream);
  box_free(b, b->d_wrec_w); box_free(b, b->d_wrec_e); box_free(b, b->d_wrec_f); box_free(b, b->d_wpart);
  b->d_wrec_w = b->d_wrec_e = b->d_wpart = nullptr;
  b->d_wrec_f = nullptr;
  b->wrec_cap = 0;
  const size_t n = (size_t)n_insert;
  const size_t nblk = (n + 2047) / 2048;
  int rc = hip_check(box_malloc(b, &b->d_wrec_w, n * sizeof(double)), "hipMalloc widom records");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_wrec_e, n * sizeof(double)), "hipMalloc widom records");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_wrec_f, n * sizeof(int)), "hipMalloc widom records");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_wpart, 6 * nblk * sizeof(double)), "hipMalloc widom records");
  if (!rc) b->wrec_cap = n;
  return rc;
}

static uint32_t spread3(uint32_t v)
{
  uint32_t r = 0;
  for (int i = 0; i < 10; ++i) r |= ((v >> i) & 1u) << (3 * i);
  return r;
}

// Cell grid of the Widom overlap screen.  Cells are at least as wide
// (perpendicular to each face, in fractional coordinates) as the largest
// overlap distance of any type pair, so every atom that can overlap a point
// lies in the 27 cells around the point's cell: testing those is a complete
// overlap test.  The table holds cap slots per cell (fixed-point x,y,z +
// type, empty slots typed ~0u).  The Widom kernel's FP64 copy of the
// framework is sorted by the Morton code of the atoms' cells, so that 64
// consecutive atoms are close in space.  Without a usable grid (fewer than
// three cells along an axis, or more than kCellCapMax atoms in a cell) the
// screen scans all atoms.
int graspa::build_screen_cells(graspa_box* b)
{
  constexpr int kCellCapMax = 16;
  box_free(b, b->d_fws); box_free(b, b->d_fws_type); box_free(b, b->d_ctab);
  b->d_fws = nullptr; b->d_fws_type = nullptr; b->d_ctab = nullptr;
  b->use_cells = 0;
  b->cell_cap = 0;
  b->cell_n[0] = b->cell_n[1] = b->cell_n[2] = 0;
  const BoxConst& bc = b->bc;
  const int n = b->n_fw, T = bc.ntypes;
  double thr_max = 0.01;
  for (int i = 0; i < T * T; ++i)
    thr_max = std::max(thr_max, overlap_r2(4.0 * b->eps[i], b->sig[i] * b->sig[i], b->shift[i], bc.overlap, bc.cut_vdw2));
  const double d = std::sqrt(thr_max) * (1.0 + 1e-3);
  const double* c = bc.cell;
  auto cross_norm = [](const double* u, const double* v) {
    const double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
    return std::sqrt(x * x + y * y + z * z);
  };
  const double w[3] = {b->volume / cross_norm(c + 3, c + 6), b->volume / cross_norm(c + 6, c), b->volume / cross_norm(c, c + 3)};
  int nc[3];
  bool grid_ok = n > 0;
  for (int a = 0; a < 3; ++a) {
    nc[a] = (int)std::min(64.0, std::floor(w[a] / d));
    if (nc[a] < 3) grid_ok = false;
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::vector<int> cell_of(n, 0);
  std::vector<uint32_t> key(n, 0);
  std::vector<uint32_t> fx(3 * (size_t)std::max(n, 1));
  for (int i = 0; i < n; ++i) fixed_from_cart(bc, &b->fw_pos[3 * i], &fx[3 * i]);
  int cap = 0;
  if (grid_ok) {
    std::vector<int> count((size_t)nc[0] * nc[1] * nc[2], 0);
    for (int i = 0; i < n; ++i) {
      uint32_t ci[3];
      for (int a = 0; a < 3; ++a) ci[a] = (uint32_t)(((uint64_t)fx[3 * i + a] * (uint64_t)nc[a]) >> 32);
      cell_of[i] = (int)(ci[0] + nc[0] * (ci[1] + nc[1] * ci[2]));
      key[i] = spread3(ci[0]) | (spread3(ci[1]) << 1) | (spread3(ci[2]) << 2);
      cap = std::max(cap, ++count[cell_of[i]]);
    }
    if (cap > kCellCapMax) grid_ok = false;
  }
  if (grid_ok) std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key[x] < key[y]; });
  const size_t m = (size_t)std::max(n, 1);
  std::vector<double> soa(4 * m, 0.0);
  std::vector<int> ty(m, 0);
  for (int k = 0; k < n; ++k) {
    const int i = order[k];
    soa[k] = b->fw_pos[3 * i];
    soa[n + k] = b->fw_pos[3 * i + 1];
    soa[2 * n + k] = b->fw_pos[3 * i + 2];
    soa[3 * n + k] = b->fw_q[i];
    ty[k] = b->fw_type[i];
  }
  int rc = hip_check(box_malloc(b, &b->d_fws, 4 * m * sizeof(double)), "hipMalloc fw");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_fws_type, m * sizeof(int)), "hipMalloc fw");
  rc = rc ? rc : hip_check(hipMemcpy(b->d_fws, soa.data(), 4 * m * sizeof(double), hipMemcpyHostToDevice), "fw upload");
  rc = rc ? rc : hip_check(hipMemcpy(b->d_fws_type, ty.data(), m * sizeof(int), hipMemcpyHostToDevice), "fw upload");
  if (rc || !grid_ok) return rc;
  const size_t ncell = (size_t)nc[0] * nc[1] * nc[2];
  std::vector<uint4> tab(ncell * cap, make_uint4(0u, 0u, 0u, ~0u));
  std::vector<int> fill(ncell, 0);
  for (int k = 0; k < n; ++k) {
    const int i = order[k];
    const int cl = cell_of[i];
    tab[(size_t)cl * cap + fill[cl]++] = make_uint4(fx[3 * i], fx[3 * i + 1], fx[3 * i + 2], (uint32_t)b->fw_type[i]);
  }
  rc = hip_check(box_malloc(b, &b->d_ctab, tab.size() * sizeof(uint4)), "hipMalloc cells");
  rc = rc ? rc : hip_check(hipMemcpy(b->d_ctab, tab.data(), tab.size() * sizeof(uint4), hipMemcpyHostToDevice), "cells upload");
  if (rc) return rc;
  b->use_cells = 1;
  b->cell_cap = cap;
  for (int a = 0; a < 3; ++a) b->cell_n[a] = nc[a];
  return GRASPA_OK;
}

#define GRASPA_CHECK_BOX(b) \
  if (!(b)) return fail(GRASPA_ERR_ARG, "null box");

extern "C" {

const char* graspa_last_error(void) { return g_last_error.c_str(); }
const char* graspa_version(void) { return "graspa_amd 0.1 (gfx950)"; }

int graspa_device_count(int* n)
{
  if (!n) return fail(GRASPA_ERR_ARG, "null");
  return hip_check(hipGetDeviceCount(n), "hipGetDeviceCount");
}

int graspa_box_create(const graspa_box_params* p, graspa_box** out)
{
  if (!p || !out) return fail(GRASPA_ERR_ARG, "null argument");
  if (p->ntypes <= 0 || p->ntypes > kMaxTypes) return fail(GRASPA_ERR_ARG, "ntypes out of range (1..32)");
  if (!p->epsilon || !p->sigma || !p->shift) return fail(GRASPA_ERR_ARG, "missing force-field tables");
  if (p->precision != GRASPA_FP64 && p->precision != GRASPA_MIXED) return fail(GRASPA_ERR_ARG, "bad precision");
  if (p->kmax[0] < 0 || p->kmax[1] < 0 || p->kmax[2] < 0 || p->kmax[0] > 31 || p->kmax[1] > 31 || p->kmax[2] > 31)
    return fail(GRASPA_ERR_ARG, "kmax out of range (0..31)");
  if (hip_check(hipSetDevice(p->device), "hipSetDevice")) return GRASPA_ERR_HIP;
  graspa_box* b = new graspa_box();
  b->device = p->device;
  b->precision = p->precision;
  if (hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate")) {
    delete b;
    return GRASPA_ERR_HIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, p->device) == hipSuccess) {
    b->n_cu = prop.multiProcessorCount;
    b->max_lds = (int)prop.sharedMemPerBlock;
  }
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, p->device) == hipSuccess &&
        v > b->max_lds)
      b->max_lds = v;
  }
  BoxConst& bc = b->bc;
  std::memcpy(bc.cell, p->cell, sizeof(bc.cell));
  // maths.cuh:28-55
  const double* x = p->cell;
  const double m11 = x[0], m21 = x[3], m31 = x[6], m12 = x[1], m22 = x[4], m32 = x[7], m13 = x[2], m23 = x[5],
               m33 = x[8];
  const double det = +m11 * (m22 * m33 - m23 * m32) - m12 * (m21 * m33 - m23 * m31) + m13 * (m21 * m32 - m22 * m31);
  if (!(std::fabs(det) > 0.0)) {
    graspa_box_destroy(b);
    return fail(GRASPA_ERR_ARG, "singular cell");
  }
  double* r = bc.inv;
  r[0] = +(m22 * m33 - m32 * m23) / det;
  r[3] = -(m21 * m33 - m31 * m23) / det;
  r[6] = +(m21 * m32 - m31 * m22) / det;
  r[1] = -(m12 * m33 - m32 * m13) / det;
  r[4] = +(m11 * m33 - m31 * m13) / det;
  r[7] = -(m11 * m32 - m31 * m12) / det;
  r[2] = +(m12 * m23 - m22 * m13) / det;
  r[5] = -(m11 * m23 - m21 * m13) / det;
  r[8] = +(m11 * m22 - m21 * m12) / det;
  b->volume = det;
  bc.cubic = (std::fabs(x[3]) + std::fabs(x[6]) + std::fabs(x[7])) > 1e-10 ? 0 : 1;
  bc.ntypes = p->ntypes;
  bc.cut_vdw2 = p->cutoff_vdw * p->cutoff_vdw;
  bc.cut_coul2 = p->cutoff_coulomb * p->cutoff_coulomb;
  bc.overlap = p->overlap_criteria;
  bc.alpha = p->alpha;
  bc.prefactor = p->prefactor;
  bc.use_charges = p->use_charges ? 1 : 0;
  for (int d = 0; d < 3; ++d) bc.kmax[d] = p->kmax[d];
  bc.nvec = (p->kmax[0] + 1) * (2 * p->kmax[1] + 1) * (2 * p->kmax[2] + 1);
  for (int i = 0; i < 9; ++i) bc.h32[i] = (float)(p->cell[i] * (1.0 / 4294967296.0));
  for (int i = 0; i < 9; ++i) bc.h64[i] = p->cell[i] * (1.0 / 4294967296.0);
  bc.cut_vdw2f = (float)bc.cut_vdw2;
  bc.cut_coul2f = (float)bc.cut_coul2;
  bc.overlapf = (float)std::min(bc.overlap, 3.0e38);
  bc.alphaf = (float)bc.alpha;
  bc.preff = (float)bc.prefactor;
  b->recip_cut = p->reciprocal_cutoff;
  b->lammps_ewald = p->lammps_ewald;
  const int T = p->ntypes;
  b->eps.assign(p->epsilon, p->epsilon + T * T);
  b->sig.assign(p->sigma, p->sigma + T * T);
  b->shift.assign(p->shift, p->shift + T * T);
  std::vector<double> ff64(4 * T * T);
  std::vector<float> ff32(4 * T * T);
  for (int i = 0; i < T * T; ++i) {
    ff64[4 * i + 0] = 4.0 * p->epsilon[i];
    ff64[4 * i + 1] = p->sigma[i] * p->sigma[i];
    ff64[4 * i + 2] = p->shift[i];
    ff64[4 * i + 3] = p->epsilon[i];
    for (int k = 0; k < 4; ++k) ff32[4 * i + k] = (float)ff64[4 * i + k];
  }
  int rc = hip_check(box_malloc(b, &b->d_ff64, ff64.size() * sizeof(double)), "hipMalloc ff");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_ff32, ff32.size() * sizeof(float)), "hipMalloc ff");
  rc = rc ? rc : hip_check(hipMemcpy(b->d_ff64, ff64.data(), ff64.size() * sizeof(double), hipMemcpyHostToDevice), "ff upload");
  rc = rc ? rc : hip_check(hipMemcpy(b->d_ff32, ff32.data(), ff32.size() * sizeof(float), hipMemcpyHostToDevice), "ff upload");
  // Ewald: per-k factor prefactor*2pi/V * (1 or 2) * exp(-k^2/4a^2)/k^2, -1 where the k-vector
  // is outside the reciprocal cutoff (Ewald_Energy_Functions.h:296-340)
  b->nvec = bc.nvec;
  b->ktemp.assign(b->nvec, -1.0);
  {
    const double PI = 3.14159265358979323846;
    const double pref = p->prefactor * (2.0 * PI / det);
    const double alpha_sq = p->alpha * p->alpha;
    const int kym = p->kmax[1], kzm = p->kmax[2];
    int nact = 0;
    for (int k = 0; k < b->nvec; ++k) {
      const int kz = k % (2 * kzm + 1) - kzm;
      const int kxy = k / (2 * kzm + 1);
      const int kx = kxy / (2 * kym + 1);
      const int ky = kxy % (2 * kym + 1) - kym;
      double ksqr = (double)(kx * kx + ky * ky + kz * kz);
      if (p->lammps_ewald) {
        const double lx = x[0], ly = x[4], lz = x[8], xy = x[3], xz = x[6], yz = x[7];
        const double ux = 2 * PI / lx, uy = 2 * PI * (-xy) / lx / ly, uz = 2 * PI * (xy * yz - ly * xz) / lx / ly / lz;
        const double vy = 2 * PI / ly, vz = 2 * PI * (-yz) / ly / lz, wz = 2 * PI / lz;
        const double kvx = kx * ux, kvy = kx * uy + ky * vy, kvz = kx * uz + ky * vz + kz * wz;
        ksqr = kvx * kvx + kvy * kvy + kvz * kvz;
      }
      if (!(ksqr > 1e-10 && ksqr < p->reciprocal_cutoff)) continue;
      const double* a = bc.inv;
      const double kvx = a[0] * 2.0 * PI * (double)kx + a[1] * 2.0 * PI * (double)ky + a[2] * 2.0 * PI * (double)kz;
      const double kvy = a[3] * 2.0 * PI * (double)kx + a[4] * 2.0 * PI * (double)ky + a[5] * 2.0 * PI * (double)kz;
      const double kvz = a[6] * 2.0 * PI * (double)kx + a[7] * 2.0 * PI * (double)ky + a[8] * 2.0 * PI * (double)kz;
      const double rksq = kvx * kvx + kvy * kvy + kvz * kvz;
      const double factor = (kx == 0) ? (1.0 * pref) : (2.0 * pref);
      b->ktemp[k] = factor * std::exp((-0.25 / alpha_sq) * rksq) / rksq;
      ++nact;
    }
    b->n_active = nact;
  }
  const size_t nv = (size_t)std::max(b->nvec, 1);
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_ktemp, nv * sizeof(double)), "hipMalloc ewald");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_sk_fw, nv * sizeof(double2)), "hipMalloc ewald");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_sk_ads, nv * sizeof(double2)), "hipMalloc ewald");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_sk_stage, nv * sizeof(double2)), "hipMalloc ewald");
  if (!rc) {
    hipMemcpy(b->d_ktemp, b->ktemp.data(), b->nvec * sizeof(double), hipMemcpyHostToDevice);
    hipMemset(b->d_sk_fw, 0, nv * sizeof(double2));
    hipMemset(b->d_sk_ads, 0, nv * sizeof(double2));
    hipMemset(b->d_sk_stage, 0, nv * sizeof(double2));
  }
  {
    // compact list of the active k-vectors for the Widom kernel's Fourier term
    std::vector<int4> kact;
    std::vector<double> tact;
    const int kzm = p->kmax[2], kym = p->kmax[1];
    for (int k = 0; k < b->nvec; ++k) {
      if (b->ktemp[k] < 0.0) continue;
      const int kz = k % (2 * kzm + 1) - kzm;
      const int kxy = k / (2 * kzm + 1);
      const int kx = kxy / (2 * kym + 1);
      const int ky = kxy % (2 * kym + 1) - kym;
      kact.push_back(make_int4(kx, ky, kz, k));
      tact.push_back(b->ktemp[k]);
    }
    const size_t na = std::max<size_t>(kact.size(), 1);
    rc = rc ? rc : hip_check(box_malloc(b, &b->d_kact, na * sizeof(int4)), "hipMalloc ewald");
    rc = rc ? rc : hip_check(box_malloc(b, &b->d_ktemp_act, na * sizeof(double)), "hipMalloc ewald");
    if (!rc && !kact.empty()) {
      rc = hip_check(hipMemcpy(b->d_kact, kact.data(), kact.size() * sizeof(int4), hipMemcpyHostToDevice), "k upload");
      rc = rc ? rc
              : hip_check(hipMemcpy(b->d_ktemp_act, tact.data(), tact.size() * sizeof(double), hipMemcpyHostToDevice),
                          "k upload");
    }
  }
  {
    // real-space Coulomb table of the fast form
    std::vector<double> ct(kCoulTabStride, 0.0);
    if (bc.use_charges && bc.alpha > 0.0 && bc.cut_coul2 > 0.0) build_coulomb_table(bc.alpha, bc.cut_coul2, ct);
    b->coul_n = (int)ct.size();
    rc = rc ? rc : hip_check(box_malloc(b, &b->d_coul, ct.size() * sizeof(double)), "hipMalloc coulomb table");
    rc = rc ? rc
            : hip_check(hipMemcpy(b->d_coul, ct.data(), ct.size() * sizeof(double), hipMemcpyHostToDevice),
                        "coulomb table upload");
  }
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_wacc, kWaccSize * sizeof(double)), "hipMalloc widom");
  if (!rc) rc = hip_check(hipMemset(b->d_wacc, 0, kWaccSize * sizeof(double)), "widom init");
  rc = rc ? rc : ensure_part(b, 4096);
  b->comps.resize(1);
  rc = rc ? rc : realloc_guests(b);
  if (!rc) {
    // empty framework arrays so kernels always get valid pointers
    rc = hip_check(box_malloc(b, &b->d_fw, 4 * sizeof(double)), "hipMalloc fw");
    rc = rc ? rc : hip_check(box_malloc(b, &b->d_fw_type, sizeof(int)), "hipMalloc fw");
    rc = rc ? rc : hip_check(box_malloc(b, &b->d_fw_fix, sizeof(uint4)), "hipMalloc fw");
    rc = rc ? rc : hip_check(box_malloc(b, &b->d_fw_fixT, sizeof(uint4)), "hipMalloc fw");
    rc = rc ? rc : build_screen_cells(b);
  }
  if (rc) {
    graspa_box_destroy(b);
    return rc;
  }
  *out = b;
  return GRASPA_OK;
}

void graspa_box_destroy(graspa_box* b)
{
  if (!b) return;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  timing_clear(b);
  widom_graph_free(b);
  box_free(b, b->d_ff64); box_free(b, b->d_ff32);
  box_free(b, b->d_fw); box_free(b, b->d_fw_type); box_free(b, b->d_fw_fix); box_free(b, b->d_fw_fixT);
  box_free(b, b->d_fws); box_free(b, b->d_fws_type); box_free(b, b->d_ctab);
  box_free(b, b->d_g); box_free(b, b->d_g_type); box_free(b, b->d_g_molid); box_free(b, b->d_g_comp); box_free(b, b->d_g_fix);
  box_free(b, b->d_ktemp); box_free(b, b->d_sk_fw); box_free(b, b->d_sk_ads); box_free(b, b->d_sk_stage);
  box_free(b, b->d_kact); box_free(b, b->d_ktemp_act); box_free(b, b->d_coul);
  free_dead_maps(b);
  for (Component& c : b->comps) box_free(b, c.d_chain);
  box_free(b, b->d_part);
  if (b->h_part) hipHostFree(b->h_part);
  box_free(b, b->d_wacc); box_free(b, b->d_wres);
  box_free(b, b->d_wrec_w); box_free(b, b->d_wrec_e); box_free(b, b->d_wrec_f); box_free(b, b->d_wpart);
  box_free(b, b->d_wmask1); box_free(b, b->d_wmask2); box_free(b, b->d_we1); box_free(b, b->d_we2);
  box_free(b, b->d_wlist1); box_free(b, b->d_wlist2); box_free(b, b->d_wstate); box_free(b, b->d_wcounts); box_free(b, b->d_wlive); box_free(b, b->d_wstat);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
}

int graspa_box_set_framework(graspa_box* b, int n, const double* xyz, const double* q, const int* type)
{
  GRASPA_CHECK_BOX(b);
  if (n < 0 || (n > 0 && (!xyz || !q || !type))) return fail(GRASPA_ERR_ARG, "bad framework arrays");
  for (int i = 0; i < n; ++i)
    if (type[i] < 0 || type[i] >= b->bc.ntypes) return fail(GRASPA_ERR_ARG, "framework type out of range");
  hipSetDevice(b->device);
  hipStreamSynchronize(b->stream);
  box_free(b, b->d_fw); box_free(b, b->d_fw_type); box_free(b, b->d_fw_fix); box_free(b, b->d_fw_fixT);
  b->d_fw = nullptr; b->d_fw_type = nullptr; b->d_fw_fix = nullptr; b->d_fw_fixT = nullptr;
  free_dead_maps(b);
  b->n_fw = n;
  b->fw_pos.assign(xyz, xyz + 3 * (size_t)n);
  b->fw_q.assign(q, q + n);
  b->fw_type.assign(type, type + n);
  const size_t m = (size_t)std::max(n, 1);
  std::vector<double> soa(4 * m, 0.0);
  std::vector<uint4> fix(m, make_uint4(0, 0, 0, 0));
  std::vector<uint4> fixT(m, make_uint4(0, 0, 0, 0));
  for (int i = 0; i < n; ++i) {
    soa[i] = xyz[3 * i];
    soa[n + i] = xyz[3 * i + 1];
    soa[2 * n + i] = xyz[3 * i + 2];
    soa[3 * n + i] = q[i];
    uint32_t f[3];
    fixed_from_cart(b->bc, &xyz[3 * i], f);
    float qf = (float)q[i];
    uint32_t qb;
    std::memcpy(&qb, &qf, 4);
    fix[i] = make_uint4(f[0], f[1], f[2], qb);
    fixT[i] = make_uint4(f[0], f[1], f[2], (uint32_t)type[i]);
  }
  int rc = hip_check(box_malloc(b, &b->d_fw, 4 * m * sizeof(double)), "hipMalloc fw");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_fw_type, m * sizeof(int)), "hipMalloc fw");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_fw_fix, m * sizeof(uint4)), "hipMalloc fw");
  rc = rc ? rc : hip_check(box_malloc(b, &b->d_fw_fixT, m * sizeof(uint4)), "hipMalloc fw");
  rc = rc ? rc : hip_check(hipMemcpy(b->d_fw_fixT, fixT.data(), m * sizeof(uint4), hipMemcpyHostToDevice), "fw upload");
  rc = rc ? rc : hip_check(hipMemcpy(b->d_fw, soa.data(), 4 * m * sizeof(double), hipMemcpyHostToDevice), "fw upload");
  if (!rc && n > 0) rc = hip_check(hipMemcpy(b->d_fw_type, type, n * sizeof(int), hipMemcpyHostToDevice), "fw upload");
  rc = rc ? rc : hip_check(hipMemcpy(b->d_fw_fix, fix.data(), m * sizeof(uint4), hipMemcpyHostToDevice), "fw upload");
  rc = rc ? rc : build_screen_cells(b);
  return rc;
}

int graspa_box_add_component(graspa_box* b, int molsize, int capacity, const double* txyz, const int* ttype,
                             const double* tq, int* comp_out)
{
  GRASPA_CHECK_BOX(b);
  if (molsize < 1 || molsize > kMaxSites / 2 || capacity < 1 || !txyz || !ttype || !tq || !comp_out)
    return fail(GRASPA_ERR_ARG, "bad component (1 <= molsize <= 32, capacity >= 1)");
  if ((int)b->comps.size() >= kMaxComps) return fail(GRASPA_ERR_CAPACITY, "at most 7 adsorbate components");
  for (int i = 0; i < molsize; ++i)
    if (ttype[i] < 0 || ttype[i] >= b->bc.ntypes) return fail(GRASPA_ERR_ARG, "template type out of range");
  hipSetDevice(b->device);
  Component c;
  c.molsize = molsize;
  c.capacity = capacity;
  c.n_mol = 0;
  c.pos.assign(3 * (size_t)capacity * molsize, 0.0);
  c.q.assign((size_t)capacity * molsize, 0.0);
  c.type.assign((size_t)capacity * molsize, 0);
  c.molid.assign((size_t)capacity * molsize, 0);
  for (int i = 0; i < molsize; ++i) {
    for (int d = 0; d < 3; ++d) c.pos[3 * i + d] = txyz[3 * i + d];
    c.q[i] = tq[i];
    c.type[i] = ttype[i];
  }
  b->comps.push_back(c);
  *comp_out = (int)b->comps.size() - 1;
  return realloc_guests(b);
}

int graspa_box_footprint(graspa_box* b, uint64_t* bytes)
{
  GRASPA_CHECK_BOX(b);
  if (!bytes) return fail(GRASPA_ERR_ARG, "null output");
  *bytes = (uint64_t)b->dev_bytes;
  return GRASPA_OK;
}

int graspa_box_set_chain(graspa_box* b, int comp, const graspa_chain* chain)
{
  GRASPA_CHECK_BOX(b);
  if (comp < 1 || comp >= (int)b->comps.size()) return fail(GRASPA_ERR_ARG, "bad component");
  Component& c = b->comps[comp];
  hipSetDevice(b->device);
  if (!chain) {
    if (c.d_chain) {
      hipStreamSynchronize(b->stream);
      box_free(b, c.d_chain);
    }
    c.d_chain = nullptr;
    return GRASPA_OK;
  }
  if (c.molsize > kChainBeads) return fail(GRASPA_ERR_UNSUPPORTED, "flexible chains of at most 4 beads");
  for (int a = 0; a < c.molsize; ++a)
    if (c.q[a] != 0.0) return fail(GRASPA_ERR_UNSUPPORTED, "flexible chains must be uncharged");
  double m[kChainDoubles] = {0.0};
  for (int i = 0; i < kChainBeads; ++i) {
    double* d = m + kChainPer * i;
    d[0] = chain->bond[i];
    d[1] = chain->theta0[i];
    d[2] = chain->k_bend[i];
    for (int t = 0; t < 4; ++t) d[3 + t] = chain->torsion[i][t];
    if (i >= 1 && i < c.molsize && !(d[0] > 0.0)) return fail(GRASPA_ERR_ARG, "chain: bond lengths must be > 0");
    if (i >= 2 && i < c.molsize && !(d[2] > 0.0 && d[1] > 0.0 && d[1] < 3.14159265358979323846))
      return fail(GRASPA_ERR_ARG, "chain: bends need k_bend > 0 and 0 < theta0 < pi");
  }
  if (!c.d_chain && hip_check(box_malloc(b, &c.d_chain, sizeof(m)), "hipMalloc chain")) return GRASPA_ERR_HIP;
  hipStreamSynchronize(b->stream);
  return hip_check(hipMemcpy(c.d_chain, m, sizeof(m), hipMemcpyHostToDevice), "chain upload");
}

int graspa_box_set_molecules(graspa_box* b, int comp, int n_mol, const double* xyz, const double* q, const int* type,
                             const int64_t* molid)
{
  GRASPA_CHECK_BOX(b);
  if (comp < 1 || comp >= (int)b->comps.size()) return fail(GRASPA_ERR_ARG, "bad component");
  Component& c = b->comps[comp];
  if (n_mol < 0 || n_mol > c.capacity) return fail(GRASPA_ERR_CAPACITY, "more molecules than capacity");
  if (n_mol > 0 && (!xyz || !q || !type || !molid)) return fail(GRASPA_ERR_ARG, "null arrays");
  hipSetDevice(b->device);
  const int na = n_mol * c.molsize;
  for (int i = 0; i < na; ++i) {
    if (type[i] < 0 || type[i] >= b->bc.ntypes) return fail(GRASPA_ERR_ARG, "type out of range");
    for (int d = 0; d < 3; ++d) c.pos[3 * i + d] = xyz[3 * i + d];
    c.q[i] = q[i];
    c.type[i] = type[i];
    c.molid[i] = molid[i];
  }
  c.n_mol = n_mol;
  if (int rc = upload_guest_atoms(b, comp, 0, std::max(n_mol, 1) * c.molsize)) return rc;
  return hip_check(hipStreamSynchronize(b->stream), "set_molecules");
}

int graspa_box_get_molecules(graspa_box* b, int comp, int* n_mol, double* xyz, double* q, int* type, int64_t* molid)
{
  GRASPA_CHECK_BOX(b);
  if (comp < 1 || comp >= (int)b->comps.size()) return fail(GRASPA_ERR_ARG, "bad component");
  const Component& c = b->comps[comp];
  if (n_mol) *n_mol = c.n_mol;
  const size_t na = (size_t)c.n_mol * c.molsize;
  if (xyz) std::memcpy(xyz, c.pos.data(), 3 * na * sizeof(double));
  if (q) std::memcpy(q, c.q.data(), na * sizeof(double));
  if (type) std::memcpy(type, c.type.data(), na * sizeof(int));
  if (molid) std::memcpy(molid, c.molid.data(), na * sizeof(int64_t));
  return GRASPA_OK;
}

int graspa_box_nvec(graspa_box* b, int* nvec, int* n_active)
{
  GRASPA_CHECK_BOX(b);
  if (nvec) *nvec = b->nvec;
  if (n_active) *n_active = b->n_active;
  return GRASPA_OK;
}

int graspa_box_init_ewald(graspa_box* b, graspa_energy* e)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  const int nb_max = (b->nvec + 3) / 4;        // one partial per workgroup of four k-vectors
  if (int rc = ensure_part(b, 3 * (size_t)nb_max + 8)) return rc;
  int nb = 0;
  if (int rc = launch_ewald_build(b, b->d_part, &nb)) return rc;
  if (hip_check(hipMemcpy(b->h_part, b->d_part, 3 * nb * sizeof(double), hipMemcpyDeviceToHost), "ewald copy"))
    return GRASPA_ERR_HIP;
  b->sk_staged = false;
  if (e) {
    std::memset(e, 0, sizeof(*e));
    for (int i = 0; i < nb; ++i) {
      e->gg_ewald += b->h_part[3 * i];
      e->hh_ewald += b->h_part[3 * i + 1];
      e->hg_ewald += b->h_part[3 * i + 2];
    }
  }
  return GRASPA_OK;
}

int graspa_box_get_structure_factors(graspa_box* b, double* sk_fw, double* sk_ads)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  hipStreamSynchronize(b->stream);
  int rc = GRASPA_OK;
  if (sk_fw) rc = hip_check(hipMemcpy(sk_fw, b->d_sk_fw, b->nvec * sizeof(double2), hipMemcpyDeviceToHost), "sk copy");
  if (!rc && sk_ads)
    rc = hip_check(hipMemcpy(sk_ads, b->d_sk_ads, b->nvec * sizeof(double2), hipMemcpyDeviceToHost), "sk copy");
  return rc;
}

int graspa_box_set_structure_factors(graspa_box* b, const double* sk_fw, const double* sk_ads)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  hipStreamSynchronize(b->stream);
  int rc = GRASPA_OK;
  if (sk_fw) rc = hip_check(hipMemcpy(b->d_sk_fw, sk_fw, b->nvec * sizeof(double2), hipMemcpyHostToDevice), "sk copy");
  if (!rc && sk_ads)
    rc = hip_check(hipMemcpy(b->d_sk_ads, sk_ads, b->nvec * sizeof(double2), hipMemcpyHostToDevice), "sk copy");
  b->sk_staged = false;
  return rc;
}

// ---- per-move energies ----

static int run_sites(graspa_box* b, int n, const double* xyz, const int* type, const double* q, int self_comp,
                     int64_t self_molid, int excl_comp, int64_t excl_molid, double* out4, uint8_t* ov)
{
  if (n < 0 || (n > 0 && (!xyz || !type || !q))) return fail(GRASPA_ERR_ARG, "bad site arrays");
  for (int i = 0; i < n; ++i)
    if (type[i] < 0 || type[i] >= b->bc.ntypes) return fail(GRASPA_ERR_ARG, "site type out of range");
  if (int rc = ensure_part(b, 5 * (size_t)kMaxSites)) return rc;
  const int sc = self_comp, ec = excl_comp;
  const int sm = (int)self_molid, em = (int)excl_molid;
  for (int off = 0; off < n; off += kMaxSites) {
    SiteBatch sb{};
    sb.n = std::min(kMaxSites, n - off);
    for (int i = 0; i < sb.n; ++i) {
      sb.x[i] = xyz[3 * (off + i)];
      sb.y[i] = xyz[3 * (off + i) + 1];
      sb.z[i] = xyz[3 * (off + i) + 2];
      sb.q[i] = q[off + i];
      sb.type[i] = type[off + i];
    }
    if (int rc = launch_site_energies(b, sb, sc, sm, ec, em, b->d_part)) return rc;
    if (hip_check(hipMemcpyAsync(b->h_part, b->d_part, 5 * sb.n * sizeof(double), hipMemcpyDeviceToHost, b->stream),
                  "site copy"))
      return GRASPA_ERR_HIP;
    if (hip_check(hipStreamSynchronize(b->stream), "site sync")) return GRASPA_ERR_HIP;
    for (int i = 0; i < sb.n; ++i) {
      for (int k = 0; k < 4; ++k) out4[4 * (off + i) + k] = b->h_part[5 * i + k];
      if (ov) ov[off + i] = b->h_part[5 * i + 4] != 0.0 ? 1 : 0;
    }
  }
  return GRASPA_OK;
}

int graspa_site_energies(graspa_box* b, int n, const double* xyz, const int* type, const double* q, int self_comp,
                         int64_t self_molid, int excl_comp, int64_t excl_molid, double* out4, uint8_t* ov)
{
  GRASPA_CHECK_BOX(b);
  if (!out4) return fail(GRASPA_ERR_ARG, "null output");
  hipSetDevice(b->device);
  return run_sites(b, n, xyz, type, q, self_comp, self_molid, excl_comp, excl_molid, out4, ov);
}

int graspa_trial_energies(graspa_box* b, int n_trials, int chain, const double* xyz, const int* type, const double* q,
                          int self_comp, int64_t self_molid, int excl_comp, int64_t excl_molid, double* out4,
                          uint8_t* ov)
{
  GRASPA_CHECK_BOX(b);
  if (n_trials < 0 || chain < 1 || !out4 || !type || !q) return fail(GRASPA_ERR_ARG, "bad trial batch");
  hipSetDevice(b->device);
  const int n = n_trials * chain;
  std::vector<int> st(n);
  std::vector<double> sq(n);
  for (int t = 0; t < n_trials; ++t)
    for (int a = 0; a < chain; ++a) {
      st[t * chain + a] = type[a];
      sq[t * chain + a] = q[a];
    }
  std::vector<double> e(4 * (size_t)n);
  std::vector<uint8_t> o(n);
  if (int rc = run_sites(b, n, xyz, st.data(), sq.data(), self_comp, self_molid, excl_comp, excl_molid, e.data(), o.data()))
    return rc;
  for (int t = 0; t < n_trials; ++t) {
    double s[4] = {0, 0, 0, 0};
    uint8_t f = 0;
    for (int a = 0; a < chain; ++a) {
      for (int k = 0; k < 4; ++k) s[k] += e[4 * (t * chain + a) + k];
      f |= o[t * chain + a];
    }
    for (int k = 0; k < 4; ++k) out4[4 * t + k] = s[k];
    if (ov) ov[t] = f;
  }
  return GRASPA_OK;
}

int graspa_single_body_delta(graspa_box* b, int comp, int64_t molid, int M, const double* old_xyz,
                             const double* new_xyz, const int* type, const double* q, graspa_energy* de, int* overlap)
{
  GRASPA_CHECK_BOX(b);
  if (M < 1 || 2 * M > kMaxSites || !old_xyz || !new_xyz || !type || !q || !de) return fail(GRASPA_ERR_ARG, "bad move");
  hipSetDevice(b->device);
  std::vector<double> xyz(6 * (size_t)M);
  std::vector<int> st(2 * M);
  std::vector<double> sq(2 * M);
  for (int i = 0; i < M; ++i) {
    for (int d = 0; d < 3; ++d) {
      xyz[3 * i + d] = new_xyz[3 * i + d];
      xyz[3 * (M + i) + d] = old_xyz[3 * i + d];
    }
    st[i] = st[M + i] = type[i];
    sq[i] = sq[M + i] = q[i];
  }
  std::vector<double> e(8 * (size_t)M);
  std::vector<uint8_t> o(2 * M);
  if (int rc = run_sites(b, 2 * M, xyz.data(), st.data(), sq.data(), comp, molid, -1, -1, e.data(), o.data())) return rc;
  std::memset(de, 0, sizeof(*de));
  int ov = 0;
  for (int i = 0; i < M; ++i) {
    de->hg_vdw += e[4 * i] - e[4 * (M + i)];
    de->hg_real += e[4 * i + 1] - e[4 * (M + i) + 1];
    de->gg_vdw += e[4 * i + 2] - e[4 * (M + i) + 2];
    de->gg_real += e[4 * i + 3] - e[4 * (M + i) + 3];
    ov |= o[i];
  }
  if (overlap) *overlap = ov;
  return GRASPA_OK;
}

int graspa_ewald_delta(graspa_box* b, int n_old, const double* old_xyz, const double* old_q, int n_new,
                       const double* new_xyz, const double* new_q, double* same, double* cross)
{
  GRASPA_CHECK_BOX(b);
  if (n_old < 0 || n_new < 0 || n_old + n_new > kMaxSites || !same || !cross) return fail(GRASPA_ERR_ARG, "bad ewald move");
  if ((n_old && (!old_xyz || !old_q)) || (n_new && (!new_xyz || !new_q))) return fail(GRASPA_ERR_ARG, "null arrays");
  hipSetDevice(b->device);
  SiteBatch sb{};
  sb.n = n_old + n_new;
  for (int i = 0; i < sb.n; ++i) {
    const double* p = i < n_old ? &old_xyz[3 * i] : &new_xyz[3 * (i - n_old)];
    sb.x[i] = p[0]; sb.y[i] = p[1]; sb.z[i] = p[2];
    sb.q[i] = i < n_old ? old_q[i] : new_q[i - n_old];
    sb.type[i] = 0;
  }
  const int nb_max = (b->nvec + kBlock - 1) / kBlock;
  if (int rc = ensure_part(b, 2 * (size_t)nb_max + 8)) return rc;
  int nb = 0;
  if (int rc = launch_ewald_delta(b, sb, n_old, b->d_part, &nb)) return rc;
  if (hip_check(hipMemcpyAsync(b->h_part, b->d_part, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, b->stream),
                "ewald copy"))
    return GRASPA_ERR_HIP;
  if (hip_check(hipStreamSynchronize(b->stream), "ewald sync")) return GRASPA_ERR_HIP;
  double s = 0.0, c = 0.0;
  for (int i = 0; i < nb; ++i) {
    s += b->h_part[2 * i];
    c += b->h_part[2 * i + 1];
  }
  *same = s;
  *cross = c;
  b->sk_staged = true;
  return GRASPA_OK;
}

int graspa_ewald_accept(graspa_box* b)
{
  GRASPA_CHECK_BOX(b);
  if (!b->sk_staged) return fail(GRASPA_ERR_STATE, "no staged structure factor");
  std::swap(b->d_sk_ads, b->d_sk_stage);
  b->sk_staged = false;
  return GRASPA_OK;
}

int graspa_commit_move(graspa_box* b, int comp, int mol, const double* xyz)
{
  GRASPA_CHECK_BOX(b);
  if (comp < 1 || comp >= (int)b->comps.size() || !xyz) return fail(GRASPA_ERR_ARG, "bad commit");
  Component& c = b->comps[comp];
  if (mol < 0 || mol >= c.n_mol) return fail(GRASPA_ERR_ARG, "molecule slot out of range");
  hipSetDevice(b->device);
  const int a0 = mol * c.molsize;
  for (int i = 0; i < c.molsize; ++i)
    for (int d = 0; d < 3; ++d) c.pos[3 * (a0 + i) + d] = xyz[3 * i + d];
  return upload_guest_atoms(b, comp, a0, c.molsize);
}

int graspa_commit_insert(graspa_box* b, int comp, const double* xyz, const double* q, const int* type, int64_t molid)
{
  GRASPA_CHECK_BOX(b);
  if (comp < 1 || comp >= (int)b->comps.size() || !xyz || !q || !type) return fail(GRASPA_ERR_ARG, "bad commit");
  Component& c = b->comps[comp];
  if (c.n_mol + 1 > c.capacity) return fail(GRASPA_ERR_CAPACITY, "Need to allocate more space");
  hipSetDevice(b->device);
  const int a0 = c.n_mol * c.molsize;
  for (int i = 0; i < c.molsize; ++i) {
    if (type[i] < 0 || type[i] >= b->bc.ntypes) return fail(GRASPA_ERR_ARG, "type out of range");
    for (int d = 0; d < 3; ++d) c.pos[3 * (a0 + i) + d] = xyz[3 * i + d];
    c.q[a0 + i] = q[i];
    c.type[a0 + i] = type[i];
    c.molid[a0 + i] = molid;
  }
  c.n_mol++;
  return upload_guest_atoms(b, comp, a0, c.molsize);
}

int graspa_commit_delete(graspa_box* b, int comp, int mol)
{
  GRASPA_CHECK_BOX(b);
  if (comp < 1 || comp >= (int)b->comps.size()) return fail(GRASPA_ERR_ARG, "bad commit");
  Component& c = b->comps[comp];
  if (mol < 0 || mol >= c.n_mol) return fail(GRASPA_ERR_ARG, "molecule slot out of range");
  hipSetDevice(b->device);
  const int M = c.molsize, a0 = mol * M, last = (c.n_mol - 1) * M;
  if (a0 != last) {
    // the last molecule moves into the hole: positions, charges and types are
    // copied, the hole keeps its own MolIDs (Update_deletion_data_Parallel)
    for (int i = 0; i < M; ++i) {
      for (int d = 0; d < 3; ++d) c.pos[3 * (a0 + i) + d] = c.pos[3 * (last + i) + d];
      c.q[a0 + i] = c.q[last + i];
      c.type[a0 + i] = c.type[last + i];
    }
  }
  c.n_mol--;
  if (a0 != last) return upload_guest_atoms(b, comp, a0, M);
  return GRASPA_OK;
}

int graspa_total_energy(graspa_box* b, graspa_energy* e)
{
  GRASPA_CHECK_BOX(b);
  if (!e) return fail(GRASPA_ERR_ARG, "null output");
  hipSetDevice(b->device);
  GuestView gv;
  guest_view(b, &gv);
  const int nb_max = (std::max(gv.total, 1) + kBlock - 1) / kBlock;
  if (int rc = ensure_part(b, 4 * (size_t)nb_max + 8)) return rc;
  int nb = 0;
  if (int rc = launch_total_pairs(b, b->d_part, &nb)) return rc;
  if (hip_check(hipMemcpyAsync(b->h_part, b->d_part, 4 * nb * sizeof(double), hipMemcpyDeviceToHost, b->stream),
                "total copy"))
    return GRASPA_ERR_HIP;
  if (hip_check(hipStreamSynchronize(b->stream), "total sync")) return GRASPA_ERR_HIP;
  std::memset(e, 0, sizeof(*e));
  for (int i = 0; i < nb; ++i) {
    e->hg_vdw += b->h_part[4 * i];
    e->hg_real += b->h_part[4 * i + 1];
    e->gg_vdw += b->h_part[4 * i + 2];
    e->gg_real += b->h_part[4 * i + 3];
  }
  // Fourier part from the running structure factors
  std::vector<double2> sf(b->nvec), sa(b->nvec);
  if (hip_check(hipMemcpy(sf.data(), b->d_sk_fw, b->nvec * sizeof(double2), hipMemcpyDeviceToHost), "sk copy") ||
      hip_check(hipMemcpy(sa.data(), b->d_sk_ads, b->nvec * sizeof(double2), hipMemcpyDeviceToHost), "sk copy"))
    return GRASPA_ERR_HIP;
  for (int k = 0; k < b->nvec; ++k) {
    const double t = b->ktemp[k];
    if (t < 0.0) continue;
    e->gg_ewald += t * (sa[k].x * sa[k].x + sa[k].y * sa[k].y);
    e->hh_ewald += t * (sf[k].x * sf[k].x + sf[k].y * sf[k].y);
    e->hg_ewald += t * (sf[k].x * sa[k].x + sf[k].y * sa[k].y) * 2.0;
  }
  return GRASPA_OK;
}

int graspa_widom_batch(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                       uint64_t seed, uint32_t stream, int64_t first_index, int n_insert, graspa_widom_result* results,
                       graspa_widom_summary* summary)
{
  GRASPA_CHECK_BOX(b);
  if (n_insert < 0) return fail(GRASPA_ERR_ARG, "n_insert < 0");
  hipSetDevice(b->device);
  graspa_widom_result* d_res = nullptr;
  if (results) {
    if ((size_t)n_insert > b->wres_cap) {
      box_free(b, b->d_wres);
      b->d_wres = nullptr;
      if (hip_check(box_malloc(b, &b->d_wres, (size_t)n_insert * sizeof(graspa_widom_result)), "hipMalloc widom results"))
        return GRASPA_ERR_HIP;
      b->wres_cap = n_insert;
    }
    d_res = b->d_wres;
  }
  int rc = launch_widom(b, comp, K, O, beta, bias, excl, tail_de, seed, stream, first_index, n_insert, d_res, summary,
                        WidomLaunch());
  if (rc) return rc;
  if (results && n_insert > 0) {
    if (hip_check(hipMemcpyAsync(results, d_res, (size_t)n_insert * sizeof(graspa_widom_result), hipMemcpyDeviceToHost,
                                 b->stream),
                  "widom results copy"))
      return GRASPA_ERR_HIP;
  }
  return hip_check(hipStreamSynchronize(b->stream), "widom sync");
}

int graspa_widom_enqueue(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                         uint64_t seed, uint32_t stream, int64_t first_index, int n_insert)
{
  GRASPA_CHECK_BOX(b);
  if (n_insert < 0) return fail(GRASPA_ERR_ARG, "n_insert < 0");
  hipSetDevice(b->device);
  return launch_widom(b, comp, K, O, beta, bias, excl, tail_de, seed, stream, first_index, n_insert, nullptr, nullptr,
                      WidomLaunch());
}

int graspa_widom_enqueue_noalloc(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl,
                                 double tail_de, uint64_t seed, uint32_t stream, int64_t first_index, int n_insert)
{
  GRASPA_CHECK_BOX(b);
  if (n_insert < 0) return fail(GRASPA_ERR_ARG, "n_insert < 0");
  hipSetDevice(b->device);
  WidomLaunch noalloc;
  noalloc.no_alloc = true;
  return launch_widom(b, comp, K, O, beta, bias, excl, tail_de, seed, stream, first_index, n_insert, nullptr, nullptr,
                      noalloc);
}

int graspa_widom_graph_create(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl,
                              double tail_de, uint64_t seed, uint32_t stream, int64_t first_index, int n_insert)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  return widom_graph_create(b, comp, K, O, beta, bias, excl, tail_de, seed, stream, first_index, n_insert);
}

int graspa_widom_graph_launch(graspa_box* b, int n_replays)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  return widom_graph_launch(b, n_replays);
}

int graspa_widom_totals(graspa_box* b, graspa_widom_summary* totals, int reset)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  double acc[kWaccSize - kWaccTotals];
  if (hip_check(hipMemcpyAsync(acc, b->d_wacc + kWaccTotals, sizeof(acc), hipMemcpyDeviceToHost, b->stream),
                "widom totals copy"))
    return GRASPA_ERR_HIP;
  if (reset && hip_check(hipMemsetAsync(b->d_wacc + kWaccTotals, 0, sizeof(acc), b->stream), "widom totals reset"))
    return GRASPA_ERR_HIP;
  if (hip_check(hipStreamSynchronize(b->stream), "widom sync")) return GRASPA_ERR_HIP;
  if (totals) {
    unsigned long long st[2];
    std::memcpy(st, acc + 8, sizeof(st));
    *totals = graspa_widom_summary{acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], (double)st[0], (double)st[1]};
  }
  return GRASPA_OK;
}

int graspa_widom_set_shape(graspa_box* b, const graspa_eval_shape* shape)
{
  GRASPA_CHECK_BOX(b);
  const graspa_eval_shape s = shape ? *shape : graspa_eval_shape{12, 0, 1};
  if ((s.waves != 4 && s.waves != 8 && s.waves != 12 && s.waves != 16) || s.per_cu < 0 || s.per_cu > 64 ||
      (s.framework_lds != 0 && s.framework_lds != 1))
    return fail(GRASPA_ERR_ARG, "widom shape: waves 4/8/12/16, 0 <= per_cu <= 64, framework_lds 0/1");
  b->ev_waves = s.waves;
  b->ev_per_cu = s.per_cu;
  b->ev_fw_lds = s.framework_lds;
  return GRASPA_OK;
}

int graspa_widom_get_shape(graspa_box* b, graspa_eval_shape* shape)
{
  GRASPA_CHECK_BOX(b);
  if (!shape) return fail(GRASPA_ERR_ARG, "null shape");
  *shape = graspa_eval_shape{b->ev_waves, b->ev_per_cu, b->ev_fw_lds};
  return GRASPA_OK;
}

int graspa_widom_autotune(graspa_box* b, int comp, int K, int O, double beta, int bias, double excl, double tail_de,
                          uint64_t seed, int n_insert, int reps, graspa_eval_shape* tried, double* ms, int cap,
                          int* n_tried)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  return widom_autotune(b, comp, K, O, beta, bias, excl, tail_de, seed, n_insert, reps, tried, ms, cap, n_tried);
}

int graspa_timing_enable(graspa_box* b, const char* family, int on)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  hipStreamSynchronize(b->stream);
  timing_clear(b);
  b->timing.on = on != 0;
  b->timing.family = family ? family : "";
  return GRASPA_OK;
}

int graspa_timing_read(graspa_box* b, double* total_ms, int64_t* launches)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  if (hip_check(hipStreamSynchronize(b->stream), "timing sync")) return GRASPA_ERR_HIP;
  double t = 0.0;
  for (size_t i = 0; i + 1 < b->timing.ev.size(); i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->timing.ev[i], b->timing.ev[i + 1]) == hipSuccess) t += ms;
  }
  if (total_ms) *total_ms = t;
  if (launches) *launches = b->timing.launches;
  return GRASPA_OK;
}

int graspa_box_sync(graspa_box* b)
{
  GRASPA_CHECK_BOX(b);
  hipSetDevice(b->device);
  return hip_check(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
}

}  // extern "C"
