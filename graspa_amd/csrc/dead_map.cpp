// dead_map.cpp -- per-site-type maps of the framework's certain-overlap volume.
//
// The box is cut into voxels of about 0.15 A (nv[a] per cell axis, in cell
// fractions).  For a probe site of type s, a voxel is marked when every point
// of it lies within the overlap distance r_ov(s, t) of some framework atom of
// type t: the farthest corner of the voxel (a voxel is a parallelepiped, so
// the farthest point is a corner) is closer than r_ov, with r_ov^2 lowered by
// 1e-5 relative.  A probe site in a marked voxel has a pair with the framework
// whose energy exceeds OverlapCriteria (or r^2 < 0.01) -- the reference flags
// it as an overlap (VDW_Coulomb.cu:1299-1300) -- so a single bit decides most
// trial positions in a dense framework; unmarked voxels are left to the cell
// screen and the exact evaluation.  Positions map to voxels through their
// 32-bit fixed-point cell fractions: voxel = (f * nv) >> 32 per axis.
//
// The second map of a site type marks the certainly free voxels: every point
// is farther than the overlap distance from every framework atom (the voxel's
// nearest point beyond r_ov, with r_ov^2 raised by 1e-4 relative), so the cell
// screen, whose thresholds lie below r_ov^2, cannot flag a trial there and the
// screen is skipped; such trials go to the exact evaluation as before.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"

namespace graspa {

void dead_map_dims(const graspa_box* b, int nv[3])
{
  const double* c = b->bc.cell;
  auto cross_norm = [](const double* u, const double* v) {
    const double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
    return std::sqrt(x * x + y * y + z * z);
  };
  const double w[3] = {b->volume / cross_norm(c + 3, c + 6), b->volume / cross_norm(c + 6, c),
                       b->volume / cross_norm(c, c + 3)};
  for (int a = 0; a < 3; ++a) nv[a] = std::max(4, std::min(1024, (int)std::ceil(w[a] / kDeadVoxel)));
}

void build_dead_map(const graspa_box* b, int site_type, std::vector<uint32_t>& bits, int nv[3],
                    std::vector<uint32_t>* free_bits)
{
  dead_map_dims(b, nv);
  const size_t nvox = (size_t)nv[0] * nv[1] * nv[2];
  bits.assign((nvox + 31) / 32, 0u);
  if (free_bits) {
    free_bits->assign((nvox + 31) / 32, ~0u);
    if (nvox & 31) free_bits->back() = (1u << (nvox & 31)) - 1u;
  }
  const BoxConst& bc = b->bc;
  const int T = bc.ntypes;
  const double* cell = bc.cell;   // columns a = cell[0..2], b = cell[3..5], c = cell[6..8]
  std::vector<double> thr(T), thr_near(T);
  for (int t = 0; t < T; ++t) {
    const int ij = t * T + site_type;
    const double r2 = overlap_r2(4.0 * b->eps[ij], b->sig[ij] * b->sig[ij], b->shift[ij], bc.overlap, bc.cut_vdw2);
    thr[t] = r2 * (1.0 - 1.0e-5);
    thr_near[t] = r2 * (1.0 + 1.0e-4);
  }
  double wdt[3];
  {
    auto cross_norm = [](const double* u, const double* v) {
      const double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
      return std::sqrt(x * x + y * y + z * z);
    };
    wdt[0] = b->volume / cross_norm(cell + 3, cell + 6);
    wdt[1] = b->volume / cross_norm(cell + 6, cell);
    wdt[2] = b->volume / cross_norm(cell, cell + 3);
  }
  const bool ortho = cell[1] == 0.0 && cell[2] == 0.0 && cell[3] == 0.0 && cell[5] == 0.0 && cell[6] == 0.0 &&
                     cell[7] == 0.0;
  // a voxel's circumradius (half its longest body diagonal), for the nearest
  // point of a skewed voxel: no point is closer than |centre| - radius
  double vrad = 0.0;
  for (int d = 0; d < 4; ++d) {
    const double g[3] = {1.0 / nv[0], (d & 1 ? -1.0 : 1.0) / nv[1], (d & 2 ? -1.0 : 1.0) / nv[2]};
    const double dx = cell[0] * g[0] + cell[3] * g[1] + cell[6] * g[2];
    const double dy = cell[1] * g[0] + cell[4] * g[1] + cell[7] * g[2];
    const double dz = cell[2] * g[0] + cell[5] * g[1] + cell[8] * g[2];
    vrad = std::max(vrad, 0.5 * std::sqrt(dx * dx + dy * dy + dz * dz));
  }
  for (int i = 0; i < b->n_fw; ++i) {
    const double t2 = thr[b->fw_type[i]], tn2 = thr_near[b->fw_type[i]];
    if (!(t2 > 0.0)) continue;
    const double r = std::sqrt(free_bits ? tn2 : t2);
    const double* p = &b->fw_pos[3 * i];
    double s[3];
    for (int a = 0; a < 3; ++a) s[a] = bc.inv[a] * p[0] + bc.inv[3 + a] * p[1] + bc.inv[6 + a] * p[2];
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      const double ext = r / wdt[a];   // the sphere lies within +-ext in cell fraction a
      lo[a] = (int)std::floor((s[a] - ext) * nv[a]);
      hi[a] = (int)std::floor((s[a] + ext) * nv[a]);
    }
    for (int iz = lo[2]; iz <= hi[2]; ++iz)
      for (int iy = lo[1]; iy <= hi[1]; ++iy)
        for (int ix = lo[0]; ix <= hi[0]; ++ix) {
          const int ii[3] = {ix, iy, iz};
          const int wx = ((ix % nv[0]) + nv[0]) % nv[0];
          const int wy = ((iy % nv[1]) + nv[1]) % nv[1];
          const int wz = ((iz % nv[2]) + nv[2]) % nv[2];
          const size_t lin = (size_t)wx + (size_t)nv[0] * ((size_t)wy + (size_t)nv[1] * wz);
          if (free_bits) {
            // nearest point of the voxel to the atom
            double near2 = 0.0;
            if (ortho) {
              for (int a = 0; a < 3; ++a) {
                const double lo_f = (double)ii[a] / nv[a], hi_f = (double)(ii[a] + 1) / nv[a];
                const double d = s[a] < lo_f ? lo_f - s[a] : (s[a] > hi_f ? s[a] - hi_f : 0.0);
                near2 += d * d * cell[4 * a] * cell[4 * a];
              }
            } else {
              double df[3];
              for (int a = 0; a < 3; ++a) df[a] = ((double)ii[a] + 0.5) / nv[a] - s[a];
              const double dx = cell[0] * df[0] + cell[3] * df[1] + cell[6] * df[2];
              const double dy = cell[1] * df[0] + cell[4] * df[1] + cell[7] * df[2];
              const double dz = cell[2] * df[0] + cell[5] * df[1] + cell[8] * df[2];
              const double dc = std::max(0.0, std::sqrt(dx * dx + dy * dy + dz * dz) - vrad);
              near2 = dc * dc;
            }
            if (near2 >= tn2) continue;
            (*free_bits)[lin >> 5] &= ~(1u << (lin & 31));
          }
          // farthest corner of the voxel from the atom
          double far2 = 0.0;
          if (ortho) {
            // axis-aligned voxel: the farthest corner takes the farther face on each axis
            for (int a = 0; a < 3; ++a) {
              const double d0 = ((double)ii[a] / nv[a] - s[a]) * cell[4 * a];
              const double d1 = ((double)(ii[a] + 1) / nv[a] - s[a]) * cell[4 * a];
              far2 += std::max(d0 * d0, d1 * d1);
            }
          } else {
            for (int corner = 0; corner < 8; ++corner) {
              double df[3];
              for (int a = 0; a < 3; ++a) df[a] = (double)(ii[a] + ((corner >> a) & 1)) / nv[a] - s[a];
              const double dx = cell[0] * df[0] + cell[3] * df[1] + cell[6] * df[2];
              const double dy = cell[1] * df[0] + cell[4] * df[1] + cell[7] * df[2];
              const double dz = cell[2] * df[0] + cell[5] * df[1] + cell[8] * df[2];
              far2 = std::max(far2, dx * dx + dy * dy + dz * dz);
              if (far2 >= t2) break;
            }
          }
          if (far2 >= t2) continue;
          bits[lin >> 5] |= 1u << (lin & 31);
        }
  }
}

// Device copy of the map of one site type, built on first use and kept until
// the framework changes.
int ensure_dead_map(graspa_box* b, int site_type, const uint32_t** out)
{
  *out = nullptr;
  if (site_type < 0 || site_type >= b->bc.ntypes || b->n_fw == 0) return GRASPA_OK;
  if ((int)b->d_dmap.size() < b->bc.ntypes) b->d_dmap.resize(b->bc.ntypes, nullptr);
  if ((int)b->d_fmap.size() < b->bc.ntypes) b->d_fmap.resize(b->bc.ntypes, nullptr);
  if (!b->d_dmap[site_type]) {
    std::vector<uint32_t> bits, fbits;
    int nv[3];
    build_dead_map(b, site_type, bits, nv, &fbits);
    uint32_t *d = nullptr, *f = nullptr;
    const size_t bytes = bits.size() * sizeof(uint32_t);
    if (hip_check(box_malloc(b, &d, bytes), "hipMalloc dead map")) return GRASPA_ERR_HIP;
    if (hip_check(box_malloc(b, &f, bytes), "hipMalloc free map")) {
      box_free(b, d);
      return GRASPA_ERR_HIP;
    }
    if (hip_check(hipMemcpy(d, bits.data(), bytes, hipMemcpyHostToDevice), "dead map upload") ||
        hip_check(hipMemcpy(f, fbits.data(), bytes, hipMemcpyHostToDevice), "free map upload")) {
      box_free(b, d);
      box_free(b, f);
      return GRASPA_ERR_HIP;
    }
    b->d_dmap[site_type] = d;
    b->d_fmap[site_type] = f;
    for (int a = 0; a < 3; ++a) b->dmap_nv[a] = nv[a];
  }
  *out = b->d_dmap[site_type];
  return GRASPA_OK;
}

void free_dead_maps(graspa_box* b)
{
  for (uint32_t*& d : b->d_dmap) {
    box_free(b, d);
    d = nullptr;
  }
  for (uint32_t*& f : b->d_fmap) {
    box_free(b, f);
    f = nullptr;
  }
}

}  // namespace graspa

using namespace graspa;

// Both host entry points: the map of the framework (dead = 0: certain
// overlap, 1: certainly free) from the box parameters.
static int host_voxel_map(const graspa_box_params* p, int n_fw, const double* xyz, const int* type, int site_type,
                          uint32_t* bits, int cap, int* nv, int which)
{
  if (!p || n_fw < 0 || (n_fw > 0 && (!xyz || !type)) || !nv || cap < 0 || p->ntypes < 1 ||
      site_type < 0 || site_type >= p->ntypes)
    return -1;
  graspa_box b;
  std::memcpy(b.bc.cell, p->cell, sizeof(b.bc.cell));
  const double* c = p->cell;
  // inverse of the cell matrix (columns a, b, c)
  const double det = c[0] * (c[4] * c[8] - c[7] * c[5]) - c[3] * (c[1] * c[8] - c[7] * c[2]) +
                     c[6] * (c[1] * c[5] - c[4] * c[2]);
  if (!(std::fabs(det) > 0.0)) return -1;
  double* inv = b.bc.inv;
  inv[0] = (c[4] * c[8] - c[5] * c[7]) / det;
  inv[1] = (c[2] * c[7] - c[1] * c[8]) / det;
  inv[2] = (c[1] * c[5] - c[2] * c[4]) / det;
  inv[3] = (c[5] * c[6] - c[3] * c[8]) / det;
  inv[4] = (c[0] * c[8] - c[2] * c[6]) / det;
  inv[5] = (c[2] * c[3] - c[0] * c[5]) / det;
  inv[6] = (c[3] * c[7] - c[4] * c[6]) / det;
  inv[7] = (c[1] * c[6] - c[0] * c[7]) / det;
  inv[8] = (c[0] * c[4] - c[1] * c[3]) / det;
  b.volume = std::fabs(det);
  b.bc.ntypes = p->ntypes;
  b.bc.overlap = p->overlap_criteria;
  b.bc.cut_vdw2 = p->cutoff_vdw * p->cutoff_vdw;
  const int T = p->ntypes;
  b.eps.assign(p->epsilon, p->epsilon + T * T);
  b.sig.assign(p->sigma, p->sigma + T * T);
  b.shift.assign(p->shift, p->shift + T * T);
  b.n_fw = n_fw;
  b.fw_pos.assign(xyz, xyz + 3 * (size_t)n_fw);
  b.fw_type.assign(type, type + n_fw);
  std::vector<uint32_t> w, f;
  build_dead_map(&b, site_type, w, nv, which ? &f : nullptr);
  if (which) w.swap(f);
  if ((size_t)cap < w.size()) return (int)w.size();
  if (bits) std::memcpy(bits, w.data(), w.size() * sizeof(uint32_t));
  return (int)w.size();
}

extern "C" int graspa_dead_map(const graspa_box_params* p, int n_fw, const double* xyz, const int* type,
                               int site_type, uint32_t* bits, int cap, int* nv)
{
  return host_voxel_map(p, n_fw, xyz, type, site_type, bits, cap, nv, 0);
}

extern "C" int graspa_free_map(const graspa_box_params* p, int n_fw, const double* xyz, const int* type,
                               int site_type, uint32_t* bits, int cap, int* nv)
{
  return host_voxel_map(p, n_fw, xyz, type, site_type, bits, cap, nv, 1);
}
