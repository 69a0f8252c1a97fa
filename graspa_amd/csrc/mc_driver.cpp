// mc_driver.cpp -- GCMC move driver over the engine's per-move entry points,
// reproducing the reference's random-number order so that the plain FP64 form
// gives the reference's accept/reject sequence (parity mode).
//
// The driver is the caller side of the hot path: RunMoves (axpy.cu:102-298),
// SingleBodyMove (mc_single_particle.h:10-314), InsertionMove / DeletionMove
// (move_struct.h:4-168) with Insertion_Body / Deletion_Body
// (mc_swap_utilities.h:3-225) and the CBMC trial functions (mc_widom.h).
// Proposals are computed on the host in FP64 and every energy comes from the
// engine; random numbers follow RandomNumber (data_struct.h:1287-1346): a pool
// of uniform triples refilled from rand(), plus direct rand() draws.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.hpp"

namespace graspa {

// glibc random_r TYPE_3 (additive feedback, r[i] = r[i-31] + r[i-3], output
// r >> 1), seeded by the Park-Miller minimal standard generator; the published
// description of the C library's rand().
class GlibcRand {
 public:
  explicit GlibcRand(unsigned seed = 1) { srand(seed); }
  void srand(unsigned seed)
  {
    int32_t r[34];
    r[0] = (int32_t)(seed == 0 ? 1u : seed);
    for (int i = 1; i < 31; ++i) {
      const int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
      int64_t w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      r[i] = (int32_t)w;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 0; i < 34; ++i) ring_[i] = (uint32_t)r[i];
    pos_ = 34;
    for (int i = 34; i < 344; ++i) next_raw();
  }
  int rand() { return (int)(next_raw() >> 1); }

 private:
  uint32_t next_raw()
  {
    const uint32_t v = ring_[(pos_ - 31) % 34] + ring_[(pos_ - 3) % 34];
    ring_[pos_ % 34] = v;
    ++pos_;
    return v;
  }
  uint32_t ring_[34];
  uint64_t pos_ = 0;
};

}  // namespace graspa

using namespace graspa;

struct graspa_mc {
  graspa_box* box = nullptr;
  graspa_mc_params p{};
  GlibcRand rng;
  std::vector<double> pool;   // 3 * pool_size
  size_t pool_off = 0;
  int64_t total_molecules = 1;
  int64_t n_pseudo[GRASPA_MAX_PSEUDO] = {0};
  // trans_tot, trans_acc, rot_tot, rot_acc, ins_tot, ins_acc, del_tot, del_acc, widom, reins_tot, reins_acc,
  // identity swaps to each species (attempted, accepted)
  int64_t cnt[8][GRASPA_N_COUNTERS] = {{0}};
  int64_t upd_tot[8][2] = {{0}}, upd_acc[8][2] = {{0}};  // since the last max-step update
  double widom_w[8] = {0}, widom_n[8] = {0};
  graspa_energy running{};
  std::vector<double> cycle_seconds;
  std::vector<double> tail, use_tail_d;
  // creation phase (InsertionMove::CreateMolecule, move_struct.h:96-106):
  // insertions accepted with the uniform 1e-100, no move counters
  int create = 0;
  std::vector<int> use_tail;
  int ntypes = 0;

  double uniform() { return (double)rng.rand() / 2147483647.0; }  // data_struct.cpp:10-15
  void pool_reset()
  {
    pool_off = 0;
    const size_t n = pool.size() / 3;
    for (size_t i = 0; i < n; ++i) {
      pool[3 * i] = uniform();
      pool[3 * i + 1] = uniform();
      pool[3 * i + 2] = uniform();
    }
    for (size_t i = n * 3; i < 1000000; ++i) (void)uniform();
  }
  void pool_check(size_t k)
  {
    if (pool_off + k >= pool.size() / 3) pool_reset();
  }
};

namespace {

double etotal(const graspa_energy& e)
{
  return e.hh_vdw + e.hg_vdw + e.gg_vdw + e.hh_real + e.hg_real + e.gg_real + e.hh_ewald + e.hg_ewald + e.gg_ewald +
         e.tail;
}
void eadd(graspa_energy& a, const graspa_energy& b)
{
  a.hh_vdw += b.hh_vdw; a.hg_vdw += b.hg_vdw; a.gg_vdw += b.gg_vdw;
  a.hh_real += b.hh_real; a.hg_real += b.hg_real; a.gg_real += b.gg_real;
  a.hh_ewald += b.hh_ewald; a.hg_ewald += b.hg_ewald; a.gg_ewald += b.gg_ewald;
  a.tail += b.tail;
}

// Rotate_Quaternions (mc_utilities.h:423-457)
void rotate_quaternion(double v[3], double u, double vv, double w)
{
  const double pi = 3.14159265358979323846;
  const double q0 = std::sqrt(1 - u) * std::sin(2 * pi * vv);
  const double q1 = std::sqrt(1 - u) * std::cos(2 * pi * vv);
  const double q2 = std::sqrt(u) * std::sin(2 * pi * w);
  const double q3 = std::sqrt(u) * std::cos(2 * pi * w);
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

// RotationAroundAxis (mc_utilities.h:459-483)
void rotate_axis(double p[3], double theta, double ax, double ay, double az)
{
  const double c = std::cos(theta), w = 1.0 - c, s = std::sin(theta);
  double rot[9];
  rot[0] = ax * ax * w + c;      rot[1] = ax * ay * w + az * s; rot[2] = ax * az * w - ay * s;
  rot[3] = ax * ay * w - az * s; rot[4] = ay * ay * w + c;      rot[5] = ay * az * w + ax * s;
  rot[6] = ax * az * w + ay * s; rot[7] = ay * az * w - ax * s; rot[8] = az * az * w + c;
  const double x = p[0] * rot[0] + p[1] * rot[1] + p[2] * rot[2];
  const double y = p[0] * rot[3] + p[1] * rot[4] + p[2] * rot[5];
  const double z = p[0] * rot[6] + p[1] * rot[7] + p[2] * rot[8];
  p[0] = x; p[1] = y; p[2] = z;
}

// SelectTrialPosition (mc_widom.h:14-39)
int select_trial(const std::vector<double>& logw, double u)
{
  double largest = logw[0];
  for (size_t i = 1; i < logw.size(); ++i) largest = std::max(largest, logw[i]);
  std::vector<double> sh(logw.size());
  double sum = 0.0;
  for (size_t i = 0; i < logw.size(); ++i) { sh[i] = std::exp(logw[i] - largest); sum += sh[i]; }
  int sel = 0;
  double cumw = sh[0];
  const double ws = u * sum;
  while (cumw < ws && sel + 1 < (int)sh.size()) cumw += sh[++sel];
  return sel;
}

struct Trials {
  std::vector<int> keep;
  std::vector<double> logw;
  std::vector<std::array<double, 4>> e;
};

// Host_sum_Widom_HGGG_SEPARATE (mc_widom.h:41-87)
void survivors(const std::vector<double>& e4, const std::vector<uint8_t>& ov, int n, double beta, int bias, Trials& t)
{
  t.keep.clear(); t.logw.clear(); t.e.clear();
  for (int i = 0; i < n; ++i) {
    if (ov[i]) continue;
    std::array<double, 4> v = {e4[4 * i], e4[4 * i + 1], e4[4 * i + 2], e4[4 * i + 3]};
    double tot = v[0] + v[2];
    if (bias) tot += v[1] + v[3];
    t.keep.push_back(i);
    t.logw.push_back(-beta * tot);
    t.e.push_back(v);
  }
}

double sum_exp(const Trials& t)
{
  double s = 0.0;
  for (double l : t.logw) s += std::exp(l);
  return s;
}

double tail_delta(const graspa_mc* mc, int comp, int sign)
{
  // TailCorrectionDifference (TailCorrection_Energy_Functions.h:36-82)
  if (!mc->p.has_tail || mc->tail.empty()) return 0.0;
  const int T = mc->ntypes;
  double e = 0.0;
  const int64_t* d = mc->p.species[comp - 1].pseudo_per_mol;
  for (int i = 0; i < T && i < GRASPA_MAX_PSEUDO; ++i) {
    const int64_t di = d[i] * sign;
    for (int j = i; j < T && j < GRASPA_MAX_PSEUDO; ++j) {
      const int64_t dj = d[j] * sign;
      const int ij = i * T + j;
      if (!mc->use_tail[ij]) continue;
      const int64_t dn = mc->n_pseudo[i] * dj + mc->n_pseudo[j] * di + di * dj;
      double de = mc->tail[ij] * (double)dn;
      if (i != j) de *= 2.0;
      e += de;
    }
  }
  return e / mc->box->volume;
}

int single_body(graspa_mc* mc, int comp, int mol, int move, graspa_trace& tr)
{
  graspa_box* b = mc->box;
  Component& c = b->comps[comp];
  const graspa_species& sp = mc->p.species[comp - 1];
  const int M = c.molsize, start = mol * M;
  const int k = (move == GRASPA_MOVE_TRANSLATION) ? 0 : 1;
  mc->cnt[comp][2 * k]++;
  mc->upd_tot[comp][k]++;
  mc->pool_check((size_t)M);
  const double* R = &mc->pool[3 * mc->pool_off];
  mc->pool_off += (size_t)M;
  std::vector<double> oldp(3 * M), newp(3 * M), q(M);
  std::vector<int> ty(M);
  for (int i = 0; i < M; ++i) {
    for (int d = 0; d < 3; ++d) oldp[3 * i + d] = c.pos[3 * (start + i) + d];
    q[i] = c.q[start + i];
    ty[i] = c.type[start + i];
  }
  if (move == GRASPA_MOVE_TRANSLATION) {
    for (int i = 0; i < M; ++i)
      for (int d = 0; d < 3; ++d) newp[3 * i + d] = oldp[3 * i + d] + sp.max_translation[d] * 2.0 * (R[d] - 0.5);
  } else {
    double ang[3];
    for (int d = 0; d < 3; ++d) ang[d] = sp.max_rotation[d] * 2.0 * (R[d] - 0.5);
    for (int i = 0; i < M; ++i) {
      double v[3] = {oldp[3 * i] - oldp[0], oldp[3 * i + 1] - oldp[1], oldp[3 * i + 2] - oldp[2]};
      rotate_axis(v, ang[0], 1.0, 0.0, 0.0);
      rotate_axis(v, ang[1], 0.0, 1.0, 0.0);
      rotate_axis(v, ang[2], 0.0, 0.0, 1.0);
      for (int d = 0; d < 3; ++d) newp[3 * i + d] = v[d] + oldp[d];
    }
  }
  graspa_energy de{};
  int overlap = 0;
  if (int rc = graspa_single_body_delta(b, comp, c.molid[start], M, oldp.data(), newp.data(), ty.data(), q.data(), &de,
                                        &overlap))
    return rc;
  double pacc = 0.0;
  double qabs = 0.0;
  for (int i = 0; i < M; ++i) qabs += std::fabs(q[i]);
  const bool charged = b->bc.use_charges && qabs > 1e-6;   // hasPartialCharge (read_data.cpp:2160-2164)
  if (!overlap) {
    if (charged) {
      double same = 0.0, cross = 0.0;
      if (int rc = graspa_ewald_delta(b, M, oldp.data(), q.data(), M, newp.data(), q.data(), &same, &cross)) return rc;
      de.gg_ewald = same;
      de.hg_ewald = 2.0 * cross;
    }
    pacc = 1.0 * std::exp(-mc->p.beta * etotal(de));
  }
  int acc = 0;
  if (!overlap) {
    const double u = mc->uniform();
    if (u < pacc) acc = 1;
  }
  if (acc) {
    if (int rc = graspa_commit_move(b, comp, mol, newp.data())) return rc;
    if (charged)
      if (int rc = graspa_ewald_accept(b)) return rc;
    mc->cnt[comp][2 * k + 1]++;
    mc->upd_acc[comp][k]++;
    eadd(mc->running, de);
  } else {
    de = graspa_energy{};
  }
  tr.accepted = acc;
  tr.overlap = overlap;
  tr.pacc = pacc;
  tr.de = de;
  return GRASPA_OK;
}

int cbmc_move(graspa_mc* mc, int comp, int mol, int move, graspa_trace& tr, int production)
{
  graspa_box* b = mc->box;
  Component& c = b->comps[comp];
  const graspa_species& sp = mc->p.species[comp - 1];
  const int M = c.molsize, chain = M - 1;
  const int K = mc->p.n_trial_pos, O = mc->p.n_trial_orient;
  const bool del = (move == GRASPA_MOVE_DELETION);
  const int64_t N = c.n_mol;
  const int64_t sel_molid = del ? (int64_t)mol : N;
  const int start = del ? mol * M : 0;
  const double beta = mc->p.beta;
  const int bias = mc->p.vdw_real_bias;
  if (move == GRASPA_MOVE_INSERTION && !mc->create) mc->cnt[comp][4]++;
  if (move == GRASPA_MOVE_DELETION) mc->cnt[comp][6]++;
  if (move == GRASPA_MOVE_WIDOM) mc->cnt[comp][8]++;
  tr.accepted = 0; tr.overlap = 0; tr.pacc = 0.0; tr.rosen = 0.0; tr.de = graspa_energy{};
  graspa_energy en{};
  Trials ts;
  auto fail_out = [&]() {
    tr.rosen = 0.0;
    if (move == GRASPA_MOVE_WIDOM && production) mc->widom_n[comp] += 1.0;
    return GRASPA_OK;
  };
  // first bead (Widom_Move_FirstBead_PARTIAL, mc_widom.h:385-507)
  mc->pool_check((size_t)K);
  std::vector<double> sp3(3 * (size_t)std::max(K, O * std::max(chain, 1)));
  for (int k = 0; k < K; ++k) {
    const double* R = &mc->pool[3 * (mc->pool_off + k)];
    if (del && k == 0) {
      for (int d = 0; d < 3; ++d) sp3[d] = c.pos[3 * start + d];
    } else {
      sp3[3 * k] = b->bc.cell[0] * R[0];
      sp3[3 * k + 1] = b->bc.cell[4] * R[1];
      sp3[3 * k + 2] = b->bc.cell[8] * R[2];
    }
  }
  mc->pool_off += (size_t)K;
  std::vector<double> e4(4 * (size_t)std::max(K, O));
  std::vector<uint8_t> ov(std::max(K, O));
  const int t0 = c.type[start];
  const double q0 = c.q[start];
  if (int rc = graspa_trial_energies(b, K, 1, sp3.data(), &t0, &q0, comp, sel_molid, -1, -1, e4.data(), ov.data()))
    return rc;
  survivors(e4, ov, K, beta, bias, ts);
  int sel = 0;
  double rb;
  if (!del) {
    if (ts.keep.empty()) return fail_out();
    sel = select_trial(ts.logw, mc->uniform());
    rb = sum_exp(ts);
    if (rb < 1e-150) return fail_out();
  } else {
    sel = 0;
    rb = sum_exp(ts);
  }
  double W = rb / (double)K;
  if (!ts.keep.empty() && !bias) W *= std::exp(-beta * (ts.e[sel][1] + ts.e[sel][3]));
  if (W <= 1e-150) return fail_out();
  const int fbt = ts.keep[sel];
  en.hg_vdw = ts.e[sel][0]; en.hg_real = ts.e[sel][1]; en.gg_vdw = ts.e[sel][2]; en.gg_real = ts.e[sel][3];
  std::vector<double> mol_pos(3 * M), mol_q(M);
  std::vector<int> mol_t(M);
  for (int d = 0; d < 3; ++d) mol_pos[d] = sp3[3 * fbt + d];
  mol_q[0] = q0;
  mol_t[0] = t0;
  if (chain > 0) {
    // orientations (Widom_Move_Chain_PARTIAL, mc_widom.h:509-614): template
    // vectors from the component's first molecule slot as it is now
    mc->pool_check((size_t)O);
    for (int o = 0; o < O; ++o) {
      const double* R = &mc->pool[3 * (mc->pool_off + o)];
      for (int a = 0; a < chain; ++a) {
        const int s = o * chain + a;
        if (del && o == 0) {
          for (int d = 0; d < 3; ++d) sp3[3 * s + d] = c.pos[3 * (start + 1 + a) + d];
        } else {
          double v[3] = {c.pos[3 * (1 + a)] - c.pos[0], c.pos[3 * (1 + a) + 1] - c.pos[1],
                         c.pos[3 * (1 + a) + 2] - c.pos[2]};
          rotate_quaternion(v, R[0], R[1], R[2]);
          for (int d = 0; d < 3; ++d) sp3[3 * s + d] = mol_pos[d] + v[d];
        }
      }
    }
    mc->pool_off += (size_t)O;
    std::vector<int> ct(chain);
    std::vector<double> cq(chain);
    for (int a = 0; a < chain; ++a) {
      ct[a] = c.type[start + 1 + a];
      cq[a] = c.q[start + 1 + a];
    }
    if (int rc = graspa_trial_energies(b, O, chain, sp3.data(), ct.data(), cq.data(), comp, sel_molid, -1, -1,
                                       e4.data(), ov.data()))
      return rc;
    survivors(e4, ov, O, beta, bias, ts);
    int osel = 0;
    double rc_sum;
    if (!del) {
      if (ts.keep.empty()) return fail_out();
      osel = select_trial(ts.logw, mc->uniform());
      rc_sum = sum_exp(ts);
      if (rc_sum < 1e-150) return fail_out();
    } else {
      osel = 0;
      rc_sum = sum_exp(ts);
      if (ts.keep.empty()) return fail_out();
    }
    double Wc = rc_sum / (double)O;
    if (!bias) Wc *= std::exp(-beta * (ts.e[osel][1] + ts.e[osel][3]));
    W *= Wc;
    if (W <= 1e-150) return fail_out();
    const int ot = ts.keep[osel];
    en.hg_vdw += ts.e[osel][0]; en.hg_real += ts.e[osel][1]; en.gg_vdw += ts.e[osel][2]; en.gg_real += ts.e[osel][3];
    for (int a = 0; a < chain; ++a) {
      for (int d = 0; d < 3; ++d) mol_pos[3 * (1 + a) + d] = sp3[3 * (ot * chain + a) + d];
      mol_q[1 + a] = cq[a];
      mol_t[1 + a] = ct[a];
    }
  }
  // GetPrefactor (mc_utilities.h:315-351)
  const double V = b->volume, Nd = (double)N;
  const double pref = del ? Nd / (beta * sp.mol_fraction * mc->p.pressure * sp.fugacity_coeff * V)
                          : beta * sp.mol_fraction * mc->p.pressure * sp.fugacity_coeff * V / (1.0 + Nd);
  double qabs = 0.0;
  for (int i = 0; i < M; ++i) qabs += std::fabs(c.q[start + i]);
  const bool charged = b->bc.use_charges && qabs > 1e-6;
  if (charged) {
    double same = 0.0, cross = 0.0;
    const double excl = sp.excl_intra + sp.excl_self;
    if (!del) {
      if (int rc = graspa_ewald_delta(b, 0, nullptr, nullptr, M, mol_pos.data(), mol_q.data(), &same, &cross)) return rc;
      same -= excl * 1.0;
      en.gg_ewald = same;
      en.hg_ewald = 2.0 * cross;
      W *= std::exp(-beta * (same + 2.0 * cross));
    } else {
      std::vector<double> op(3 * M), oq(M);
      for (int i = 0; i < M; ++i) {
        for (int d = 0; d < 3; ++d) op[3 * i + d] = c.pos[3 * (start + i) + d];
        oq[i] = c.q[start + i];
      }
      if (int rc = graspa_ewald_delta(b, M, op.data(), oq.data(), 0, nullptr, nullptr, &same, &cross)) return rc;
      same -= excl * (0.0 - 1.0);
      W /= std::exp(-beta * (same + 2.0 * cross));
      en.gg_ewald = -1.0 * same;
      en.hg_ewald = -1.0 * (2.0 * cross);
    }
  }
  if (mc->p.has_tail) {
    const double te = tail_delta(mc, comp, del ? -1 : 1);
    if (!del) { W *= std::exp(-beta * te); en.tail = te; }
    else { W /= std::exp(-beta * te); en.tail = -te; }
  }
  tr.rosen = W;
  if (move == GRASPA_MOVE_WIDOM) {
    if (production) { mc->widom_w[comp] += W; mc->widom_n[comp] += 1.0; }
    tr.de = en;
    return GRASPA_OK;
  }
  double pacc = del ? pref * sp.ideal_rosen / W : pref * W / sp.ideal_rosen;
  if (!del && W <= 1e-150) pacc = 0.0;
  tr.pacc = pacc;
  const double u = mc->create ? 1e-100 : mc->uniform();
  if (!(u < pacc)) return GRASPA_OK;
  tr.accepted = 1;
  if (!del) {
    if (int rc = graspa_commit_insert(b, comp, mol_pos.data(), mol_q.data(), mol_t.data(), N)) return rc;
    mc->total_molecules++;
    for (int t = 0; t < GRASPA_MAX_PSEUDO; ++t) mc->n_pseudo[t] += sp.pseudo_per_mol[t];
    if (!mc->create) mc->cnt[comp][5]++;
  } else {
    if (int rc = graspa_commit_delete(b, comp, mol)) return rc;
    mc->total_molecules--;
    for (int t = 0; t < GRASPA_MAX_PSEUDO; ++t) mc->n_pseudo[t] -= sp.pseudo_per_mol[t];
    mc->cnt[comp][7]++;
    en.hg_vdw = -en.hg_vdw; en.hg_real = -en.hg_real; en.gg_vdw = -en.gg_vdw; en.gg_real = -en.gg_real;
    en.gg_ewald = -en.gg_ewald; en.hg_ewald = -en.hg_ewald; en.tail = -en.tail;
  }
  if (charged)
    if (int rc = graspa_ewald_accept(b)) return rc;
  tr.de = en;
  eadd(mc->running, en);
  return GRASPA_OK;
}

// CBMC reinsertion: ReinsertionMove::Run (move_struct.h:170-407).  Growth of
// a new configuration as an insertion (K random first beads, O orientations
// of the slot-0 template about the selected bead, energies without the
// molecule itself); retrace of the old one with one first-bead trial (its own
// position, whose weight is added to the growth's other first-bead weights,
// CBMC_FirstBead_Finish mc_widom.h:365-366) and O orientations (trial 0 = its
// own); W_new times the Ewald factor of moving the molecule; Pacc = W_new /
// W_old.  Pool draws: K triples, selection, O triples, selection, one triple
// (retrace first bead, unused), O triples, Metropolis.
int reinsertion_move(graspa_mc* mc, int comp, int mol, graspa_trace& tr)
{
  graspa_box* b = mc->box;
  Component& c = b->comps[comp];
  const int M = c.molsize, chain = M - 1;
  const int K = mc->p.n_trial_pos, O = mc->p.n_trial_orient;
  const int start = mol * M;
  const int64_t molid = c.molid[start];
  const double beta = mc->p.beta;
  const int bias = mc->p.vdw_real_bias;
  mc->cnt[comp][9]++;
  tr.accepted = 0; tr.overlap = 0; tr.pacc = 0.0; tr.rosen = 0.0; tr.de = graspa_energy{};
  std::vector<double> sp3(3 * (size_t)std::max(K, O * std::max(chain, 1)));
  std::vector<double> e4(4 * (size_t)std::max(K, O));
  std::vector<uint8_t> ov(std::max(K, O));
  std::vector<double> oldp(3 * M), newp(3 * M), q(M);
  std::vector<int> ct(std::max(chain, 1));
  std::vector<double> cq(std::max(chain, 1));
  for (int i = 0; i < M; ++i) {
    for (int d = 0; d < 3; ++d) oldp[3 * i + d] = c.pos[3 * (start + i) + d];
    q[i] = c.q[start + i];
  }
  for (int a = 0; a < chain; ++a) {
    ct[a] = c.type[start + 1 + a];
    cq[a] = c.q[start + 1 + a];
  }
  const int t0 = c.type[start];
  const double q0 = c.q[start];
  Trials ts;
  double en_new[4] = {0, 0, 0, 0}, en_old[4] = {0, 0, 0, 0};
  // growth of the new configuration
  mc->pool_check((size_t)K);
  for (int k = 0; k < K; ++k) {
    const double* R = &mc->pool[3 * (mc->pool_off + k)];
    for (int d = 0; d < 3; ++d) sp3[3 * k + d] = b->bc.cell[4 * d] * R[d];
  }
  mc->pool_off += (size_t)K;
  if (int rc = graspa_trial_energies(b, K, 1, sp3.data(), &t0, &q0, comp, molid, -1, -1, e4.data(), ov.data()))
    return rc;
  survivors(e4, ov, K, beta, bias, ts);
  if (ts.keep.empty()) return GRASPA_OK;
  const int sel = select_trial(ts.logw, mc->uniform());
  const double rb = sum_exp(ts);
  if (rb < 1e-150) return GRASPA_OK;
  const double stored_r = rb - std::exp(ts.logw[sel]);
  double Wn = rb / (double)K;
  if (!bias) Wn *= std::exp(-beta * (ts.e[sel][1] + ts.e[sel][3]));
  if (Wn <= 1e-150) return GRASPA_OK;
  for (int k = 0; k < 4; ++k) en_new[k] = ts.e[sel][k];
  for (int d = 0; d < 3; ++d) newp[d] = sp3[3 * ts.keep[sel] + d];
  if (chain > 0) {
    mc->pool_check((size_t)O);
    for (int o = 0; o < O; ++o) {
      const double* R = &mc->pool[3 * (mc->pool_off + o)];
      for (int a = 0; a < chain; ++a) {
        double v[3] = {c.pos[3 * (1 + a)] - c.pos[0], c.pos[3 * (1 + a) + 1] - c.pos[1],
                       c.pos[3 * (1 + a) + 2] - c.pos[2]};
        rotate_quaternion(v, R[0], R[1], R[2]);
        for (int d = 0; d < 3; ++d) sp3[3 * (o * chain + a) + d] = newp[d] + v[d];
      }
    }
    mc->pool_off += (size_t)O;
    if (int rc = graspa_trial_energies(b, O, chain, sp3.data(), ct.data(), cq.data(), comp, molid, -1, -1, e4.data(),
                                       ov.data()))
      return rc;
    survivors(e4, ov, O, beta, bias, ts);
    if (ts.keep.empty()) return GRASPA_OK;
    const int osel = select_trial(ts.logw, mc->uniform());
    const double rc_sum = sum_exp(ts);
    if (rc_sum < 1e-150) return GRASPA_OK;
    double Wc = rc_sum / (double)O;
    if (!bias) Wc *= std::exp(-beta * (ts.e[osel][1] + ts.e[osel][3]));
    Wn *= Wc;
    if (Wn <= 1e-150) return GRASPA_OK;
    for (int k = 0; k < 4; ++k) en_new[k] += ts.e[osel][k];
    const int ot = ts.keep[osel];
    for (int a = 0; a < chain; ++a)
      for (int d = 0; d < 3; ++d) newp[3 * (1 + a) + d] = sp3[3 * (ot * chain + a) + d];
  }
  // retrace of the old configuration
  mc->pool_check(1);
  mc->pool_off += 1;
  if (int rc = graspa_trial_energies(b, 1, 1, oldp.data(), &t0, &q0, comp, molid, -1, -1, e4.data(), ov.data()))
    return rc;
  survivors(e4, ov, 1, beta, bias, ts);
  double Wo = (sum_exp(ts) + stored_r) / (double)K;
  if (!ts.keep.empty()) {
    if (!bias) Wo *= std::exp(-beta * (ts.e[0][1] + ts.e[0][3]));
    for (int k = 0; k < 4; ++k) en_old[k] = ts.e[0][k];
  }
  if (chain > 0) {
    mc->pool_check((size_t)O);
    for (int o = 0; o < O; ++o) {
      const double* R = &mc->pool[3 * (mc->pool_off + o)];
      for (int a = 0; a < chain; ++a) {
        const int s = o * chain + a;
        if (o == 0) {
          for (int d = 0; d < 3; ++d) sp3[3 * s + d] = oldp[3 * (1 + a) + d];
        } else {
          double v[3] = {c.pos[3 * (1 + a)] - c.pos[0], c.pos[3 * (1 + a) + 1] - c.pos[1],
                         c.pos[3 * (1 + a) + 2] - c.pos[2]};
          rotate_quaternion(v, R[0], R[1], R[2]);
          for (int d = 0; d < 3; ++d) sp3[3 * s + d] = oldp[d] + v[d];
        }
      }
    }
    mc->pool_off += (size_t)O;
    if (int rc = graspa_trial_energies(b, O, chain, sp3.data(), ct.data(), cq.data(), comp, molid, -1, -1, e4.data(),
                                       ov.data()))
      return rc;
    survivors(e4, ov, O, beta, bias, ts);
    double Wc = sum_exp(ts) / (double)O;
    if (!ts.keep.empty()) {    // SelectedTrial = 0 of the survivors (mc_widom.h:588-604)
      if (!bias) Wc *= std::exp(-beta * (ts.e[0][1] + ts.e[0][3]));
      for (int k = 0; k < 4; ++k) en_old[k] += ts.e[0][k];
    }
    Wo *= Wc;
  }
  graspa_energy de{};
  de.hg_vdw = en_new[0] - en_old[0];
  de.hg_real = en_new[1] - en_old[1];
  de.gg_vdw = en_new[2] - en_old[2];
  de.gg_real = en_new[3] - en_old[3];
  double qabs = 0.0;
  for (int i = 0; i < M; ++i) qabs += std::fabs(q[i]);
  const bool charged = b->bc.use_charges && qabs > 1e-6;
  if (charged) {
    double same = 0.0, cross = 0.0;
    if (int rc = graspa_ewald_delta(b, M, oldp.data(), q.data(), M, newp.data(), q.data(), &same, &cross)) return rc;
    de.gg_ewald = same;
    de.hg_ewald = 2.0 * cross;
    Wn *= std::exp(-beta * (same + 2.0 * cross));
  }
  // ReinsertionMove::Acceptance (move_struct.h:323-358)
  const double u = mc->uniform();
  const double pacc = Wn / Wo;
  tr.rosen = Wn;
  tr.pacc = pacc;
  if (Wn <= 1e-150 || !(u < pacc)) return GRASPA_OK;
  if (int rc = graspa_commit_move(b, comp, mol, newp.data())) return rc;
  if (charged)
    if (int rc = graspa_ewald_accept(b)) return rc;
  mc->cnt[comp][10]++;
  tr.accepted = 1;
  tr.de = de;
  eadd(mc->running, de);
  return GRASPA_OK;
}

// Identity swap: IdentitySwapMove (mc_swap_moves.h:199-426).  The species
// pair is drawn two uniforms per try until the OLD species holds a molecule,
// then the OLD molecule.  The NEW molecule is grown on the OLD molecule's
// first bead (one trial whose weight is not divided by K,
// copy_firstbead_to_new :178-181) with O orientations of the NEW template,
// energies without the OLD molecule (ExcludeList); the OLD molecule is
// retraced with its own first bead and O orientations (trial 0 its own).
// W_new carries the Ewald change of the OLD sites out and the NEW sites in
// (GPU_EwaldDifference_IdentitySwap, Ewald_Energy_Functions.h:582-632) and
// the tail change (TailCorrectionIdentitySwap,
// TailCorrection_Energy_Functions.h:85-108); Pacc = pref_ins(NEW)
// pref_del(OLD) (W_new / W_id,NEW) / (W_old / W_id,OLD).  Pool draws: the
// pair, the molecule, one triple (NEW first bead, unused), O triples and a
// selection, one triple (OLD first bead, unused), O triples, Metropolis.
int identity_swap_move(graspa_mc* mc, graspa_trace& tr)
{
  graspa_box* b = mc->box;
  const int ncomp = (int)b->comps.size();
  const int K = mc->p.n_trial_pos, O = mc->p.n_trial_orient;
  const double beta = mc->p.beta;
  const int bias = mc->p.vdw_real_bias;
  tr.accepted = 0; tr.overlap = 0; tr.pacc = 0.0; tr.rosen = 0.0; tr.de = graspa_energy{};
  if (mc->total_molecules - 1 == 0) return GRASPA_OK;      // no adsorbates: no draws
  int oc = 0, nc = 0;
  int64_t nold = 0;
  while (oc == 0 || oc >= ncomp || nc == 0 || nc >= ncomp || nold == 0) {
    oc = (int)(mc->uniform() * (double)(ncomp - 1)) + 1;
    nc = (int)(mc->uniform() * (double)(ncomp - 1)) + 1;
    nold = oc < ncomp ? b->comps[oc].n_mol : 0;
  }
  Component& co = b->comps[oc];
  Component& cn = b->comps[nc];
  const int oldmol = (int)(mc->uniform() * (double)co.n_mol);
  const int Mo = co.molsize, Mn = cn.molsize, so = oldmol * Mo;
  const int64_t nnew = cn.n_mol;
  tr.comp = oc;
  tr.mol = oldmol;
  mc->cnt[oc][GRASPA_CNT_IDSW_TOT + nc - 1]++;
  const int nmax = std::max(K, O * std::max(std::max(Mo, Mn), 1));
  std::vector<double> sp3(3 * (size_t)nmax), e4(4 * (size_t)nmax);
  std::vector<uint8_t> ov(nmax);
  std::vector<double> oldp(3 * Mo), oldq(Mo), newp(3 * Mn), newq(Mn);
  std::vector<int> oldt(Mo), newt(Mn);
  for (int i = 0; i < Mo; ++i) {
    for (int d = 0; d < 3; ++d) oldp[3 * i + d] = co.pos[3 * (so + i) + d];
    oldq[i] = co.q[so + i];
    oldt[i] = co.type[so + i];
  }
  for (int i = 0; i < Mn; ++i) {
    newq[i] = cn.q[i];
    newt[i] = cn.type[i];
  }
  Trials ts;
  double en_new[4] = {0, 0, 0, 0}, en_old[4] = {0, 0, 0, 0};
  // NEW first bead at the OLD molecule's first bead
  mc->pool_check(1);
  mc->pool_off += 1;
  if (int rc = graspa_trial_energies(b, 1, 1, oldp.data(), &newt[0], &newq[0], nc, nnew, oc, oldmol, e4.data(),
                                     ov.data()))
    return rc;
  survivors(e4, ov, 1, beta, bias, ts);
  if (ts.keep.empty()) return GRASPA_OK;
  double Wn = sum_exp(ts);
  if (Wn < 1e-150) return GRASPA_OK;
  if (!bias) Wn *= std::exp(-beta * (ts.e[0][1] + ts.e[0][3]));
  if (Wn <= 1e-150) return GRASPA_OK;
  for (int k = 0; k < 4; ++k) en_new[k] = ts.e[0][k];
  for (int d = 0; d < 3; ++d) newp[d] = oldp[d];
  // NEW chain: O orientations of the NEW template about it
  if (Mn > 1) {
    const int chain = Mn - 1;
    mc->pool_check((size_t)O);
    for (int o = 0; o < O; ++o) {
      const double* R = &mc->pool[3 * (mc->pool_off + o)];
      for (int a = 0; a < chain; ++a) {
        double v[3] = {cn.pos[3 * (1 + a)] - cn.pos[0], cn.pos[3 * (1 + a) + 1] - cn.pos[1],
                       cn.pos[3 * (1 + a) + 2] - cn.pos[2]};
        rotate_quaternion(v, R[0], R[1], R[2]);
        for (int d = 0; d < 3; ++d) sp3[3 * (o * chain + a) + d] = newp[d] + v[d];
      }
    }
    mc->pool_off += (size_t)O;
    if (int rc = graspa_trial_energies(b, O, chain, sp3.data(), &newt[1], &newq[1], nc, nnew, oc, oldmol, e4.data(),
                                       ov.data()))
      return rc;
    survivors(e4, ov, O, beta, bias, ts);
    if (ts.keep.empty()) return GRASPA_OK;
    const int osel = select_trial(ts.logw, mc->uniform());
    const double rc_sum = sum_exp(ts);
    if (rc_sum < 1e-150) return GRASPA_OK;
    double Wc = rc_sum / (double)O;
    if (!bias) Wc *= std::exp(-beta * (ts.e[osel][1] + ts.e[osel][3]));
    Wn *= Wc;
    if (Wn <= 1e-150) return GRASPA_OK;
    for (int k = 0; k < 4; ++k) en_new[k] += ts.e[osel][k];
    const int ot = ts.keep[osel];
    for (int a = 0; a < chain; ++a)
      for (int d = 0; d < 3; ++d) newp[3 * (1 + a) + d] = sp3[3 * (ot * chain + a) + d];
  }
  // retrace of the OLD molecule: its own first bead
  mc->pool_check(1);
  mc->pool_off += 1;
  if (int rc = graspa_trial_energies(b, 1, 1, oldp.data(), &oldt[0], &oldq[0], oc, oldmol, -1, -1, e4.data(),
                                     ov.data()))
    return rc;
  survivors(e4, ov, 1, beta, bias, ts);
  double Wo = sum_exp(ts);
  if (!ts.keep.empty()) {
    if (!bias) Wo *= std::exp(-beta * (ts.e[0][1] + ts.e[0][3]));
    for (int k = 0; k < 4; ++k) en_old[k] = ts.e[0][k];
  }
  if (Mo > 1) {
    const int chain = Mo - 1;
    mc->pool_check((size_t)O);
    for (int o = 0; o < O; ++o) {
      const double* R = &mc->pool[3 * (mc->pool_off + o)];
      for (int a = 0; a < chain; ++a) {
        const int s = o * chain + a;
        if (o == 0) {
          for (int d = 0; d < 3; ++d) sp3[3 * s + d] = oldp[3 * (1 + a) + d];
        } else {
          double v[3] = {co.pos[3 * (1 + a)] - co.pos[0], co.pos[3 * (1 + a) + 1] - co.pos[1],
                         co.pos[3 * (1 + a) + 2] - co.pos[2]};
          rotate_quaternion(v, R[0], R[1], R[2]);
          for (int d = 0; d < 3; ++d) sp3[3 * s + d] = oldp[d] + v[d];
        }
      }
    }
    mc->pool_off += (size_t)O;
    if (int rc = graspa_trial_energies(b, O, chain, sp3.data(), &oldt[1], &oldq[1], oc, oldmol, -1, -1, e4.data(),
                                       ov.data()))
      return rc;
    survivors(e4, ov, O, beta, bias, ts);
    double Wc = sum_exp(ts) / (double)O;
    if (!ts.keep.empty()) {
      if (!bias) Wc *= std::exp(-beta * (ts.e[0][1] + ts.e[0][3]));
      for (int k = 0; k < 4; ++k) en_old[k] += ts.e[0][k];
    }
    Wo *= Wc;
  }
  graspa_energy de{};
  de.hg_vdw = en_new[0] - en_old[0];
  de.hg_real = en_new[1] - en_old[1];
  de.gg_vdw = en_new[2] - en_old[2];
  de.gg_real = en_new[3] - en_old[3];
  // Ewald: the OLD sites out, the NEW sites in, charged species only
  auto has_q = [](const std::vector<double>& q) {
    double a = 0.0;
    for (double v : q) a += std::fabs(v);
    return a > 1e-6;                 // hasPartialCharge (read_data.cpp:2160-2164)
  };
  const bool qo = b->bc.use_charges && has_q(oldq), qn = b->bc.use_charges && has_q(newq);
  const bool ewald = qo || qn;
  if (ewald) {
    double same = 0.0, cross = 0.0;
    if (int rc = graspa_ewald_delta(b, qo ? Mo : 0, oldp.data(), oldq.data(), qn ? Mn : 0, newp.data(), newq.data(),
                                    &same, &cross))
      return rc;
    const graspa_species& sn = mc->p.species[nc - 1];
    const graspa_species& so_ = mc->p.species[oc - 1];
    if (qn) same -= (sn.excl_intra + sn.excl_self) * 1.0;
    if (qo) same -= (so_.excl_intra + so_.excl_self) * -1.0;
    de.gg_ewald = same;
    de.hg_ewald = 2.0 * cross;
    Wn *= std::exp(-beta * (same + 2.0 * cross));
  }
  // tail, with the pseudo-atom change NEW - OLD
  if (mc->p.has_tail && !mc->tail.empty()) {
    const int T = mc->ntypes;
    const int64_t* pn = mc->p.species[nc - 1].pseudo_per_mol;
    const int64_t* po = mc->p.species[oc - 1].pseudo_per_mol;
    double e = 0.0;
    for (int i = 0; i < T && i < GRASPA_MAX_PSEUDO; ++i) {
      const int64_t di = pn[i] - po[i];
      for (int j = i; j < T && j < GRASPA_MAX_PSEUDO; ++j) {
        const int64_t dj = pn[j] - po[j];
        const int ij = i * T + j;
        if (!mc->use_tail[ij]) continue;
        const int64_t dn = mc->n_pseudo[i] * dj + mc->n_pseudo[j] * di + di * dj;
        double d = mc->tail[ij] * (double)dn;
        if (i != j) d *= 2.0;
        e += d;
      }
    }
    de.tail = e / b->volume;
  }
  Wn *= std::exp(-beta * de.tail);
  // GetPrefactor (mc_utilities.h:315-345) of the NEW insertion and the OLD deletion
  const graspa_species& sn = mc->p.species[nc - 1];
  const graspa_species& sq = mc->p.species[oc - 1];
  const double V = b->volume, P = mc->p.pressure;
  double pref = beta * sn.mol_fraction * P * sn.fugacity_coeff * V / (1.0 + (double)cn.n_mol);
  pref *= (double)co.n_mol / (beta * sq.mol_fraction * P * sq.fugacity_coeff * V);
  const double pacc = pref * (Wn / sn.ideal_rosen) / (Wo / sq.ideal_rosen);
  tr.rosen = Wn;
  tr.pacc = pacc;
  const double u = mc->uniform();
  if (!(u < pacc)) return GRASPA_OK;
  if (nc != oc) {
    if (cn.n_mol + 1 > cn.capacity) return fail(GRASPA_ERR_CAPACITY, "Need to allocate more space");
    if (int rc = graspa_commit_delete(b, oc, oldmol)) return rc;
    if (int rc = graspa_commit_insert(b, nc, newp.data(), newq.data(), newt.data(), nnew)) return rc;
    for (int t = 0; t < GRASPA_MAX_PSEUDO; ++t) mc->n_pseudo[t] += sn.pseudo_per_mol[t] - sq.pseudo_per_mol[t];
  } else {
    if (int rc = graspa_commit_move(b, oc, oldmol, newp.data())) return rc;
  }
  if (ewald)
    if (int rc = graspa_ewald_accept(b)) return rc;
  mc->cnt[oc][GRASPA_CNT_IDSW_ACC + nc - 1]++;
  tr.accepted = 1;
  tr.de = de;
  eadd(mc->running, de);
  return GRASPA_OK;
}

// Update_Max_Translation / Update_Max_Rotation (mc_utilities.h:612-665)
void update_max(double v[3], int64_t& acc, int64_t& tot, double hi)
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

}  // namespace

extern "C" {

int graspa_glibc_rand_sequence(unsigned seed, int n, int* out)
{
  if (n < 0 || (n > 0 && !out)) return fail(GRASPA_ERR_ARG, "bad arguments");
  GlibcRand r(seed);
  for (int i = 0; i < n; ++i) out[i] = r.rand();
  return GRASPA_OK;
}

int graspa_mc_create(graspa_box* box, const graspa_mc_params* p, graspa_mc** out)
{
  if (!box || !p || !out) return fail(GRASPA_ERR_ARG, "null argument");
  if (p->n_species < 1 || p->n_species > 7 || p->n_species + 1 != (int)box->comps.size())
    return fail(GRASPA_ERR_ARG, "n_species must match the box's adsorbate components");
  if (p->n_trial_pos < 1 || p->n_trial_orient < 1) return fail(GRASPA_ERR_ARG, "trial counts must be >= 1");
  if (p->pool_size < 16) return fail(GRASPA_ERR_ARG, "pool_size too small");
  for (size_t c = 1; c < box->comps.size(); ++c)
    if (box->comps[c].d_chain)
      return fail(GRASPA_ERR_UNSUPPORTED, "flexible chains run in the batched driver only (graspa_batch_*)");
  graspa_mc* mc = new graspa_mc();
  mc->box = box;
  mc->p = *p;
  mc->ntypes = box->bc.ntypes;
  if (p->has_tail) {
    if (!p->tail || !p->use_tail) {
      delete mc;
      return fail(GRASPA_ERR_ARG, "tail tables required when has_tail");
    }
    const int T = mc->ntypes;
    mc->tail.assign(p->tail, p->tail + T * T);
    mc->use_tail.assign(p->use_tail, p->use_tail + T * T);
  }
  for (int t = 0; t < GRASPA_MAX_PSEUDO; ++t) mc->n_pseudo[t] = p->n_pseudo[t];
  mc->total_molecules = 1;
  for (size_t c = 1; c < box->comps.size(); ++c) mc->total_molecules += box->comps[c].n_mol;
  // RandomNumber::Setup (data_struct.h:1337-1345): srand, fill, then the
  // {2.3, 4.5, 6.7} written into pool[0] by Aaccess_device_random
  mc->rng.srand(p->seed);
  mc->pool.assign(3 * (size_t)p->pool_size, 0.0);
  mc->pool_reset();
  mc->pool[0] = 2.3;
  mc->pool[1] = 4.5;
  mc->pool[2] = 6.7;
  *out = mc;
  return GRASPA_OK;
}

void graspa_mc_destroy(graspa_mc* mc) { delete mc; }

int graspa_mc_cycles(graspa_mc* mc, int first_cycle, int n_cycles, int production, graspa_trace* trace, int max_trace,
                     int64_t* n_steps)
{
  if (!mc || n_cycles < 0) return fail(GRASPA_ERR_ARG, "bad arguments");
  graspa_box* b = mc->box;
  const int ncomp = (int)b->comps.size();
  int64_t nt = 0;
  mc->cycle_seconds.clear();
  for (int cyc = first_cycle; cyc < first_cycle + n_cycles; ++cyc) {
    const auto t0 = std::chrono::steady_clock::now();
    // Determine_Number_Of_Steps (axpy.cu:580-591)
    int64_t steps = 20;
    if (steps < mc->total_molecules) steps = mc->total_molecules;
    if (mc->p.use_max_step && steps > mc->p.max_step_per_cycle) steps = mc->p.max_step_per_cycle;
    for (int64_t s = 0; s < steps; ++s) {
      // Select_Box_Component_Molecule (axpy.cu:74-101)
      int comp = 0;
      while (!(comp >= 1 && mc->p.species[comp - 1].p_swap > 1e-10)) comp = (int)(mc->uniform() * (double)ncomp);
      const int mol = (int)(mc->uniform() * (double)b->comps[comp].n_mol);
      const double r = mc->uniform();
      const graspa_species& sp = mc->p.species[comp - 1];
      graspa_trace tr{};
      tr.cycle = cyc; tr.step = (int)s; tr.comp = comp; tr.mol = mol; tr.move = GRASPA_MOVE_NONE;
      int rc = GRASPA_OK;
      if (r < sp.p_translation) {
        tr.move = GRASPA_MOVE_TRANSLATION;
        if (b->comps[comp].n_mol > 0) rc = single_body(mc, comp, mol, GRASPA_MOVE_TRANSLATION, tr);
      } else if (r < sp.p_rotation) {
        tr.move = GRASPA_MOVE_ROTATION;
        if (b->comps[comp].n_mol > 0) rc = single_body(mc, comp, mol, GRASPA_MOVE_ROTATION, tr);
      } else if (r < sp.p_widom) {
        tr.move = GRASPA_MOVE_WIDOM;
        rc = cbmc_move(mc, comp, mol, GRASPA_MOVE_WIDOM, tr, production);
      } else if (r < sp.p_reinsertion) {
        tr.move = GRASPA_MOVE_REINSERTION;
        if (b->comps[comp].n_mol > 0) rc = reinsertion_move(mc, comp, mol, tr);
      } else if (r < sp.p_identity) {
        tr.move = GRASPA_MOVE_IDENTITY_SWAP;
        rc = identity_swap_move(mc, tr);
      } else if (r < sp.p_swap) {
        if (mc->uniform() < 0.5) {
          tr.move = GRASPA_MOVE_INSERTION;
          rc = cbmc_move(mc, comp, mol, GRASPA_MOVE_INSERTION, tr, production);
        } else {
          tr.move = GRASPA_MOVE_DELETION;
          if (b->comps[comp].n_mol > 0) rc = cbmc_move(mc, comp, mol, GRASPA_MOVE_DELETION, tr, production);
        }
      }
      if (rc) return rc;
      if (trace && nt < max_trace) trace[nt] = tr;
      ++nt;
    }
    // GatherStatisticsDuringSimulation (axpy.cu:378-400)
    if (cyc % 500 == 0) {
      for (int c = 1; c < ncomp; ++c) {
        graspa_species& sp = mc->p.species[c - 1];
        if (mc->upd_tot[c][0] > 0) update_max(sp.max_translation, mc->upd_acc[c][0], mc->upd_tot[c][0], 5.0);
        if (mc->upd_tot[c][1] > 0) update_max(sp.max_rotation, mc->upd_acc[c][1], mc->upd_tot[c][1], 3.14);
      }
    }
    const auto t1 = std::chrono::steady_clock::now();
    mc->cycle_seconds.push_back(std::chrono::duration<double>(t1 - t0).count());
  }
  if (n_steps) *n_steps = nt;
  return GRASPA_OK;
}

int graspa_mc_create_molecules(graspa_mc* mc, int comp, int n, int64_t max_attempts, int* n_created,
                               graspa_energy* de)
{
  if (!mc || n < 0 || max_attempts < 0) return fail(GRASPA_ERR_ARG, "bad arguments");
  if (comp < 1 || comp >= (int)mc->box->comps.size()) return fail(GRASPA_ERR_ARG, "component out of range");
  int created = 0;
  graspa_energy sum{};
  mc->create = 1;
  int rc = GRASPA_OK;
  for (int64_t a = 0; a < max_attempts && created < n; ++a) {
    graspa_trace tr{};
    const int mol = created > 0 ? created - 1 : 0;     // TempVal.molecule (axpy.cu:349)
    rc = cbmc_move(mc, comp, mol, GRASPA_MOVE_INSERTION, tr, 0);
    if (rc) break;
    if (tr.accepted) {
      ++created;
      eadd(sum, tr.de);
    }
  }
  mc->create = 0;
  if (n_created) *n_created = created;
  if (de) *de = sum;
  if (rc) return rc;
  if (created < n) return fail(GRASPA_ERR_STATE, "Bad Insertions When Creating Molecules!");
  return GRASPA_OK;
}

int graspa_mc_stats(graspa_mc* mc, graspa_energy* running, int* n_mol, int64_t* counters, double* widom_sum_w,
                    double* widom_n)
{
  if (!mc) return fail(GRASPA_ERR_ARG, "null mc");
  const int ncomp = (int)mc->box->comps.size();
  if (running) *running = mc->running;
  for (int c = 1; c < ncomp; ++c) {
    if (n_mol) n_mol[c - 1] = mc->box->comps[c].n_mol;
    if (counters)
      for (int k = 0; k < GRASPA_N_COUNTERS; ++k) counters[GRASPA_N_COUNTERS * (c - 1) + k] = mc->cnt[c][k];
    if (widom_sum_w) widom_sum_w[c - 1] = mc->widom_w[c];
    if (widom_n) widom_n[c - 1] = mc->widom_n[c];
  }
  return GRASPA_OK;
}

int graspa_mc_cycle_times(graspa_mc* mc, double* seconds, int max_n)
{
  if (!mc) return fail(GRASPA_ERR_ARG, "null mc");
  const int n = std::min<int>(max_n, (int)mc->cycle_seconds.size());
  for (int i = 0; i < n; ++i) seconds[i] = mc->cycle_seconds[i];
  return n;
}

}  // extern "C"
