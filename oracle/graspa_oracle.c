/*
 * graspa_oracle.c -- FP64 CPU restatement of the gRASPA per-move energy path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See graspa_oracle.h for what pins
 * each part.  Citations are to the gRASPA source tree (src_clean/).
 */
#include "graspa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* maths.cuh:28-55 (matrix_determinant, inverse_matrix); read_data.cpp:2021-2025 */
void or_box_finalize(or_box* b)
{
  const double* x = b->cell;
  double m11 = x[0], m21 = x[3], m31 = x[6];
  double m12 = x[1], m22 = x[4], m32 = x[7];
  double m13 = x[2], m23 = x[5], m33 = x[8];
  double det = +m11 * (m22 * m33 - m23 * m32) - m12 * (m21 * m33 - m23 * m31) + m13 * (m21 * m32 - m22 * m31);
  double* r = b->inv_cell;
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
  b->cubic = (fabs(x[3]) + fabs(x[6]) + fabs(x[7])) > 1e-10 ? 0 : 1;
}

/* maths.cuh:427-450: minimum image, rounding half away from zero. */
void or_pbc(double* p, const double* cell, const double* inv, int cubic)
{
  if (cubic) {
    p[0] -= (double)(int)(p[0] * inv[0] + ((p[0] >= 0.0) ? 0.5 : -0.5)) * cell[0];
    p[1] -= (double)(int)(p[1] * inv[4] + ((p[1] >= 0.0) ? 0.5 : -0.5)) * cell[4];
    p[2] -= (double)(int)(p[2] * inv[8] + ((p[2] >= 0.0) ? 0.5 : -0.5)) * cell[8];
  } else {
    double s0 = inv[0] * p[0] + inv[3] * p[1] + inv[6] * p[2];
    double s1 = inv[1] * p[0] + inv[4] * p[1] + inv[7] * p[2];
    double s2 = inv[2] * p[0] + inv[5] * p[1] + inv[8] * p[2];
    s0 -= (double)(int)(s0 + ((s0 >= 0.0) ? 0.5 : -0.5));
    s1 -= (double)(int)(s1 + ((s1 >= 0.0) ? 0.5 : -0.5));
    s2 -= (double)(int)(s2 + ((s2 >= 0.0) ? 0.5 : -0.5));
    p[0] = cell[0] * s0 + cell[3] * s1 + cell[6] * s2;
    p[1] = cell[1] * s0 + cell[4] * s1 + cell[7] * s2;
    p[2] = cell[2] * s0 + cell[5] * s1 + cell[8] * s2;
  }
}

/* maths.cuh:452-494, 12-6 branch with soft-core lambda scaling. */
void or_vdw(const or_box* b, int ta, int tb, double rr, double scaling, double out[2])
{
  int row = ta * b->ntypes + tb;
  double arg1 = 4.0 * b->eps[row];
  double arg2 = b->sig[row] * b->sig[row];
  double arg3 = b->shift[row];
  double temp = rr / arg2;
  double temp3 = temp * temp * temp;
  double rri3 = 1.0 / (temp3 + 0.5 * (1.0 - scaling) * (1.0 - scaling));
  double rri6 = rri3 * rri3;
  double term = arg1 * (rri3 * (rri3 - 1.0)) - arg3;
  double dlambda = scaling * arg1 * (rri6 * (2.0 * rri3 - 1.0));
  out[0] = scaling * term;
  out[1] = scaling < 1.0 ? term + (1.0 - scaling) * dlambda : 0.0;
}

/* maths.cuh:496-500 */
double or_coulomb_real(const or_box* b, double qa, double qb, double r, double scaling)
{
  double term = qa * qb * erfc(b->alpha * r);
  return b->prefactor * scaling * term / r;
}

int or_nvec(const or_box* b)
{
  return (b->kmax[0] + 1) * (2 * b->kmax[1] + 1) * (2 * b->kmax[2] + 1);
}

/* ------------------------------------------------------------------------ */
/* Totals                                                                    */
/* ------------------------------------------------------------------------ */

/* VDW_Coulomb.cu:34-226 (VDWReal_Total_CPU): every ordered pair of atoms that
 * do not share (component, molecule), half-weighted, split HH/HG/GG. */
void or_total_vdw_real(const or_box* b, const or_atoms* comps, int ncomp, or_energy* e)
{
  double vdw[3] = {0, 0, 0}, real[3] = {0, 0, 0};
  for (int ci = 0; ci < ncomp; ci++) {
    const or_atoms* A = &comps[ci];
    int na = A->n_mol * A->molsize;
    for (int i = 0; i < na; i++) {
      for (int cj = 0; cj < ncomp; cj++) {
        int it = (ci < 1 || cj < 1) ? ((ci >= 1 || cj >= 1) ? 1 : 0) : 2;
        const or_atoms* B = &comps[cj];
        int nb = B->n_mol * B->molsize;
        for (int j = 0; j < nb; j++) {
          if (A->molid[i] == B->molid[j] && ci == cj) continue;
          double d[3] = {A->pos[3 * i] - B->pos[3 * j], A->pos[3 * i + 1] - B->pos[3 * j + 1],
                         A->pos[3 * i + 2] - B->pos[3 * j + 2]};
          or_pbc(d, b->cell, b->inv_cell, b->cubic);
          double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          if (rr < b->cut_vdw2) {
            double res[2];
            or_vdw(b, A->type[i], B->type[j], rr, 1.0, res);
            vdw[it] += 0.5 * res[0];
          }
          if (!b->no_charges && rr < b->cut_coul2) {
            double r = sqrt(rr);
            real[it] += 0.5 * or_coulomb_real(b, A->charge[i], B->charge[j], r, 1.0);
          }
        }
      }
    }
  }
  memset(e, 0, sizeof(*e));
  e->hh_vdw = vdw[0]; e->hg_vdw = vdw[1]; e->gg_vdw = vdw[2];
  e->hh_real = real[0]; e->hg_real = real[1]; e->gg_real = real[2];
}

typedef struct { double re, im; } cplx;
static inline cplx cmul(cplx a, cplx b)
{
  cplx c = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return c;
}

/* e^{i k s} for k = 0..kmax along one axis for one atom, by the complex
 * recurrence of ewald_preparation.h:60-87 / Ewald_Energy_Functions.h:96-104. */
static void eik_axis(double s, int kmax, cplx* out)
{
  out[0].re = 1.0; out[0].im = 0.0;
  if (kmax >= 1) { out[1].re = cos(s); out[1].im = sin(s); }
  for (int k = 2; k <= kmax; k++) out[k] = cmul(out[k - 1], out[1]);
}

static void frac2pi(const or_box* b, const double* p, double s[3])
{
  /* matrix_multiply_by_vector(InverseCell, pos, s); s *= 2 pi  (maths.cuh:133-138) */
  const double* a = b->inv_cell;
  s[0] = a[0] * p[0] + a[3] * p[1] + a[6] * p[2];
  s[1] = a[1] * p[0] + a[4] * p[1] + a[7] * p[2];
  s[2] = a[2] * p[0] + a[5] * p[1] + a[8] * p[2];
  s[0] *= 2 * M_PI; s[1] *= 2 * M_PI; s[2] *= 2 * M_PI;
}

/* Squared |k| used for the reciprocal cutoff test: integer kx^2+ky^2+kz^2, or
 * the LAMMPS-style Cartesian |k|^2 (Ewald_Energy_Functions.h:302-321). */
static double ksqr_cut(const or_box* b, int kx, int ky, int kz)
{
  double ksqr = (double)(kx * kx + ky * ky + kz * kz);
  if (b->lammps_ewald) {
    double lx = b->cell[0], ly = b->cell[4], lz = b->cell[8];
    double xy = b->cell[3], xz = b->cell[6], yz = b->cell[7];
    double ux = 2 * M_PI / lx, uy = 2 * M_PI * (-xy) / lx / ly;
    double uz = 2 * M_PI * (xy * yz - ly * xz) / lx / ly / lz;
    double vy = 2 * M_PI / ly, vz = 2 * M_PI * (-yz) / ly / lz, wz = 2 * M_PI / lz;
    double kvx = kx * ux, kvy = kx * uy + ky * vy, kvz = kx * uz + ky * vz + kz * wz;
    ksqr = kvx * kvx + kvy * kvy + kvz * kvz;
  }
  return ksqr;
}

/* |k|^2 with k = 2 pi (kx ax + ky ay + kz az) (ewald_preparation.h:16-18, 91-96, 175-180) */
static double rksq_of(const or_box* b, int kx, int ky, int kz)
{
  const double* a = b->inv_cell;
  double kvx = a[0] * 2.0 * M_PI * (double)kx + a[1] * 2.0 * M_PI * (double)ky + a[2] * 2.0 * M_PI * (double)kz;
  double kvy = a[3] * 2.0 * M_PI * (double)kx + a[4] * 2.0 * M_PI * (double)ky + a[5] * 2.0 * M_PI * (double)kz;
  double kvz = a[6] * 2.0 * M_PI * (double)kx + a[7] * 2.0 * M_PI * (double)ky + a[8] * 2.0 * M_PI * (double)kz;
  return kvx * kvx + kvy * kvy + kvz * kvz;
}

/* ewald_preparation.h:5-259 (Ewald_Total).  All atoms of all components
 * enter S(k); framework atoms (component 0) also enter S_fw.  The stored
 * adsorbate structure factor is S_total - S_fw. */
void or_ewald_total(const or_box* b, const or_atoms* comps, int ncomp,
                    or_energy* fourier, double* sk_fw, double* sk_ads,
                    double* self_per_comp, double* intra_per_comp, int* n_active)
{
  int kxm = b->kmax[0], kym = b->kmax[1], kzm = b->kmax[2];
  int natoms = 0;
  for (int c = 0; c < ncomp; c++) natoms += comps[c].n_mol * comps[c].molsize;
  cplx* ex = (cplx*)malloc(sizeof(cplx) * (size_t)(natoms ? natoms : 1) * (kxm + 1));
  cplx* ey = (cplx*)malloc(sizeof(cplx) * (size_t)(natoms ? natoms : 1) * (kym + 1));
  cplx* ez = (cplx*)malloc(sizeof(cplx) * (size_t)(natoms ? natoms : 1) * (kzm + 1));
  double* qs = (double*)malloc(sizeof(double) * (size_t)(natoms ? natoms : 1));
  int* isfw = (int*)malloc(sizeof(int) * (size_t)(natoms ? natoms : 1));
  cplx tmp[64];
  int count = 0;
  for (int c = 0; c < ncomp; c++) {
    int na = comps[c].n_mol * comps[c].molsize;
    for (int i = 0; i < na; i++) {
      double s[3];
      frac2pi(b, &comps[c].pos[3 * i], s);
      eik_axis(s[0], kxm, tmp); for (int k = 0; k <= kxm; k++) ex[count + k * natoms] = tmp[k];
      eik_axis(s[1], kym, tmp); for (int k = 0; k <= kym; k++) ey[count + k * natoms] = tmp[k];
      eik_axis(s[2], kzm, tmp); for (int k = 0; k <= kzm; k++) ez[count + k * natoms] = tmp[k];
      qs[count] = comps[c].charge[i];  /* scaleCoul = 1 for whole molecules */
      isfw[count] = (c < 1);
      count++;
    }
  }
  double alpha_sq = b->alpha * b->alpha;
  double prefactor = b->prefactor * (2.0 * M_PI / b->volume);
  double e_gg = 0, e_hh = 0, e_hg = 0;
  int active = 0;
  /* k-vector index (kx, ky, kz) -> (kx (2 kym + 1) + ky + kym) (2 kzm + 1) + kz + kzm, the
   * reference's loop order; rows of fixed (kx, ky) are independent (OpenMP over
   * rows, partial energies per row summed in row order afterwards) */
  const int nky = 2 * kym + 1, nkz = 2 * kzm + 1, nrows = (kxm + 1) * nky;
  double* part = (double*)calloc((size_t)4 * nrows, sizeof(double));
#pragma omp parallel
  {
  cplx* exy = (cplx*)malloc(sizeof(cplx) * (size_t)(natoms ? natoms : 1));
#pragma omp for schedule(dynamic, 1)
  for (int row = 0; row < nrows; row++) {
    const int kx = row / nky, ky = row % nky - kym;
    double factor = (kx == 0) ? (1.0 * prefactor) : (2.0 * prefactor);
    for (int i = 0; i < natoms; i++) {
      cplx t = ey[i + natoms * abs(ky)];
      if (ky < 0) t.im = -t.im;
      exy[i] = cmul(ex[i + natoms * kx], t);
    }
    for (int kz = -kzm; kz <= kzm; kz++) {
      const int nvec = row * nkz + kz + kzm;
      double ksqr = ksqr_cut(b, kx, ky, kz);
      cplx ads = {0, 0}, fw = {0, 0};
      if (ksqr > 1e-10 && ksqr < b->recip_cut) {
        for (int i = 0; i < natoms; i++) {
          cplx t = ez[i + natoms * abs(kz)];
          if (kz < 0) t.im = -t.im;
          cplx p = cmul(exy[i], t);
          ads.re += qs[i] * p.re; ads.im += qs[i] * p.im;
          if (isfw[i]) { fw.re += qs[i] * p.re; fw.im += qs[i] * p.im; }
        }
        double rksq = rksq_of(b, kx, ky, kz);
        double temp = factor * exp((-0.25 / alpha_sq) * rksq) / rksq;
        ads.re -= fw.re; ads.im -= fw.im;
        part[4 * row] += temp * (ads.re * ads.re + ads.im * ads.im);
        part[4 * row + 1] += temp * (fw.re * fw.re + fw.im * fw.im);
        part[4 * row + 2] += temp * (fw.re * ads.re + fw.im * ads.im) * 2.0;
        part[4 * row + 3] += 1.0;
      }
      if (sk_fw) { sk_fw[2 * nvec] = fw.re; sk_fw[2 * nvec + 1] = fw.im; }
      if (sk_ads) { sk_ads[2 * nvec] = ads.re; sk_ads[2 * nvec + 1] = ads.im; }
    }
  }
  free(exy);
  }
  for (int row = 0; row < nrows; row++) {
    e_gg += part[4 * row];
    e_hh += part[4 * row + 1];
    e_hg += part[4 * row + 2];
    active += (int)part[4 * row + 3];
  }
  free(part);
  memset(fourier, 0, sizeof(*fourier));
  fourier->gg_ewald = e_gg; fourier->hh_ewald = e_hh; fourier->hg_ewald = e_hg;
  if (n_active) *n_active = active;
  /* self energy (ewald_preparation.h:172-186) and intra-molecular exclusion
   * (ewald_preparation.h:188-220), per component */
  double pself = b->prefactor * b->alpha / sqrt(M_PI);
  for (int c = 0; c < ncomp; c++) {
    double se = 0.0, ie = 0.0;
    int na = comps[c].n_mol * comps[c].molsize;
    for (int i = 0; i < na; i++) se += pself * comps[c].charge[i] * comps[c].charge[i];
    int M = comps[c].molsize;
    /* pairs (i, j > i) of each molecule; the framework is one molecule of
     * all its atoms, so this is O(N^2) there: OpenMP over i, a partial sum
     * per i, summed in i order (the same result for any thread count) */
    double* pi = (double*)calloc((size_t)(M > 0 ? M : 1), sizeof(double));
    for (int m = 0; m < comps[c].n_mol && M > 0; m++) {
      const int a0 = m * M;
#pragma omp parallel for schedule(dynamic, 64)
      for (int i = a0; i < a0 + M - 1; i++) {
        double acc = 0.0;
        for (int j = i + 1; j < a0 + M; j++) {
          double d[3] = {comps[c].pos[3 * i] - comps[c].pos[3 * j], comps[c].pos[3 * i + 1] - comps[c].pos[3 * j + 1],
                         comps[c].pos[3 * i + 2] - comps[c].pos[3 * j + 2]};
          or_pbc(d, b->cell, b->inv_cell, b->cubic);
          double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          acc += b->prefactor * comps[c].charge[i] * comps[c].charge[j] * erf(b->alpha * r) / r;
        }
        pi[i - a0] = acc;
      }
      for (int i = 0; i < M - 1; i++) ie += pi[i];
    }
    free(pi);
    if (self_per_comp) self_per_comp[c] = se;
    if (intra_per_comp) intra_per_comp[c] = ie;
  }
  free(ex); free(ey); free(ez); free(qs); free(isfw);
}

/* ewald_preparation.h:261-298, 351-366: intra and self exclusion of one rigid
 * molecule, taken from the component's first molecule slot. */
void or_exclusion_rigid(const or_box* b, const or_atoms* comp, double* intra, double* self)
{
  int M = comp->molsize;
  double ie = 0.0, se = 0.0;
  for (int i = 0; i < M - 1; i++) {
    for (int j = i + 1; j < M; j++) {
      double d[3] = {comp->pos[3 * i] - comp->pos[3 * j], comp->pos[3 * i + 1] - comp->pos[3 * j + 1],
                     comp->pos[3 * i + 2] - comp->pos[3 * j + 2]};
      or_pbc(d, b->cell, b->inv_cell, b->cubic);
      double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      ie += b->prefactor * comp->charge[i] * comp->charge[j] * erf(b->alpha * r) / r;
    }
  }
  double pself = b->prefactor * b->alpha / sqrt(M_PI);
  for (int i = 0; i < M; i++) se += pself * comp->charge[i] * comp->charge[i];
  *intra = ie;
  *self = se;
}

/* TailCorrection_Energy_Functions.h:3-21 */
double or_tail_total(const or_box* b, const int64_t* n)
{
  double e = 0.0;
  int T = b->ntypes;
  for (int i = 0; i < T; i++)
    for (int j = i; j < T; j++) {
      int ij = i * T + j;
      if (!b->use_tail[ij]) continue;
      double pe = b->tail[ij] * (double)(n[i] * n[j]);
      if (i != j) pe *= 2.0;
      e += pe;
    }
  return e / b->volume;
}

/* TailCorrection_Energy_Functions.h:36-82 */
double or_tail_delta(const or_box* b, const int64_t* n, const int64_t* d, int sign)
{
  double e = 0.0;
  int T = b->ntypes;
  for (int i = 0; i < T; i++) {
    int64_t di = d[i] * sign;
    for (int j = i; j < T; j++) {
      int64_t dj = d[j] * sign;
      int ij = i * T + j;
      if (!b->use_tail[ij]) continue;
      int64_t dn = n[i] * dj + n[j] * di + di * dj;
      double de = b->tail[ij] * (double)dn;
      if (i != j) de *= 2.0;
      e += de;
    }
  }
  return e / b->volume;
}

/* ------------------------------------------------------------------------ */
/* Per-move primitives                                                       */
/* ------------------------------------------------------------------------ */

/* Pair sums of probe sites against the system: the semantics of
 * Calculate_Multiple_Trial_Energy_VDWReal (VDW_Coulomb.cu:1183-1352) and the
 * NEW/OLD halves of Calculate_Single_Body_Energy_VDWReal (VDW_Coulomb.cu:626-841):
 * own-molecule and ExcludeList atoms skipped, overlap = E_pair > OverlapCriteria
 * or r^2 < 0.01 inside the VDW cutoff. */
void or_site_energies(const or_box* b, const or_atoms* comps, int ncomp,
                      int n_sites, const double* sp, const int* st, const double* sq,
                      int self_comp, int64_t self_molid, int excl_comp, int64_t excl_molid,
                      double* out4, uint8_t* overlap)
{
  for (int s = 0; s < n_sites; s++) {
    double hv = 0, hr = 0, gv = 0, gr = 0;
    int ov = 0;
    for (int c = 0; c < ncomp; c++) {
      const or_atoms* A = &comps[c];
      int na = A->n_mol * A->molsize;
      for (int j = 0; j < na; j++) {
        if (c == excl_comp && A->molid[j] == excl_molid) continue;
        if (c == self_comp && A->molid[j] == self_molid) continue;
        double d[3] = {A->pos[3 * j] - sp[3 * s], A->pos[3 * j + 1] - sp[3 * s + 1], A->pos[3 * j + 2] - sp[3 * s + 2]};
        or_pbc(d, b->cell, b->inv_cell, b->cubic);
        double rr = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        double ev = 0.0, er = 0.0;
        if (rr < b->cut_vdw2) {
          double res[2];
          or_vdw(b, A->type[j], st[s], rr, 1.0, res);
          if (res[0] > b->overlap) ov = 1;
          if (rr < 0.01) ov = 1;
          ev = res[0];
        }
        if (!b->no_charges && rr < b->cut_coul2) {
          double r = sqrt(rr);
          er = or_coulomb_real(b, A->charge[j], sq[s], r, 1.0);
        }
        if (c < 1) { hv += ev; hr += er; } else { gv += ev; gr += er; }
      }
    }
    out4[4 * s + 0] = hv; out4[4 * s + 1] = hr; out4[4 * s + 2] = gv; out4[4 * s + 3] = gr;
    overlap[s] = (uint8_t)ov;
  }
}

/* Fourier_Ewald_Diff (Ewald_Energy_Functions.h:280-397) for one move, both the
 * same-type half (which also produces the updated S(k)) and the cross half. */
void or_ewald_delta(const or_box* b, const double* sk_same, const double* sk_cross,
                    int n_old, const double* old_pos, const double* old_q,
                    int n_new, const double* new_pos, const double* new_q,
                    double* same, double* cross, double* sk_new)
{
  int kxm = b->kmax[0], kym = b->kmax[1], kzm = b->kmax[2];
  int n = n_old + n_new;
  cplx* ex = (cplx*)malloc(sizeof(cplx) * (size_t)(n ? n : 1) * (kxm + 1));
  cplx* ey = (cplx*)malloc(sizeof(cplx) * (size_t)(n ? n : 1) * (kym + 1));
  cplx* ez = (cplx*)malloc(sizeof(cplx) * (size_t)(n ? n : 1) * (kzm + 1));
  double* q = (double*)malloc(sizeof(double) * (size_t)(n ? n : 1));
  cplx tmp[64];
  for (int i = 0; i < n; i++) {
    const double* p = (i < n_old) ? &old_pos[3 * i] : &new_pos[3 * (i - n_old)];
    q[i] = (i < n_old) ? old_q[i] : new_q[i - n_old];
    double s[3];
    frac2pi(b, p, s);
    eik_axis(s[0], kxm, tmp); for (int k = 0; k <= kxm; k++) ex[i + k * n] = tmp[k];
    eik_axis(s[1], kym, tmp); for (int k = 0; k <= kym; k++) ey[i + k * n] = tmp[k];
    eik_axis(s[2], kzm, tmp); for (int k = 0; k <= kzm; k++) ez[i + k * n] = tmp[k];
  }
  double alpha_sq = b->alpha * b->alpha;
  double prefactor = b->prefactor * (2.0 * M_PI / b->volume);
  int nvec = or_nvec(b);
  double es = 0.0, ec = 0.0;
  for (int kxyz = 0; kxyz < nvec; kxyz++) {
    int kz = kxyz % (2 * kzm + 1) - kzm;
    int kxy = kxyz / (2 * kzm + 1);
    int kx = kxy / (2 * kym + 1);
    int ky = kxy % (2 * kym + 1) - kym;
    double ksqr = ksqr_cut(b, kx, ky, kz);
    if (sk_new) { sk_new[2 * kxyz] = sk_same[2 * kxyz]; sk_new[2 * kxyz + 1] = sk_same[2 * kxyz + 1]; }
    if (!(ksqr > 1e-10 && ksqr < b->recip_cut)) continue;
    double factor = (kx == 0) ? (1.0 * prefactor) : (2.0 * prefactor);
    cplx co = {0, 0}, cn = {0, 0};
    for (int i = 0; i < n; i++) {
      cplx t = ey[i + n * abs(ky)];
      if (ky < 0) t.im = -t.im;
      cplx xy = cmul(ex[i + n * kx], t);
      t = ez[i + n * abs(kz)];
      if (kz < 0) t.im = -t.im;
      cplx p = cmul(xy, t);
      if (i < n_old) { co.re += q[i] * p.re; co.im += q[i] * p.im; }
      else { cn.re += q[i] * p.re; cn.im += q[i] * p.im; }
    }
    double rksq = rksq_of(b, kx, ky, kz);
    double temp = factor * exp((-0.25 / alpha_sq) * rksq) / rksq;
    cplx ov = {sk_same[2 * kxyz], sk_same[2 * kxyz + 1]};
    cplx nv = {ov.re + cn.re - co.re, ov.im + cn.im - co.im};
    es += temp * (nv.re * nv.re + nv.im * nv.im);
    es -= temp * (ov.re * ov.re + ov.im * ov.im);
    if (sk_new) { sk_new[2 * kxyz] = nv.re; sk_new[2 * kxyz + 1] = nv.im; }
    cplx cv = {sk_cross[2 * kxyz], sk_cross[2 * kxyz + 1]};
    ec += temp * (cv.re * (cn.re - co.re) + cv.im * (cn.im - co.im));
  }
  *same = es;
  *cross = ec;
  free(ex); free(ey); free(ez); free(q);
}

/* ------------------------------------------------------------------------ */
/* Random numbers                                                            */
/* ------------------------------------------------------------------------ */

/* data_struct.cpp:10-15: the reference draws with the C library's rand(). */
double or_rand_uniform(void) { return (double)rand() / RAND_MAX; }
void or_srand(unsigned seed) { srand(seed); }

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11.  Round function and Weyl key schedule as published. */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform double in (0,1) for draw `draw` of work item `index` in stream
 * `stream` of job seed `seed`: 53 bits from two Philox output words. */
double or_philox_uniform(uint64_t seed, uint64_t stream, uint64_t index, uint32_t draw)
{
  uint32_t ctr[4] = {(uint32_t)index, (uint32_t)(index >> 32), draw >> 1, (uint32_t)stream};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  or_philox4x32_10(ctr, key, o);
  uint32_t a = (draw & 1u) ? o[2] : o[0];
  uint32_t c = (draw & 1u) ? o[3] : o[1];
  uint64_t x = (((uint64_t)a << 32) | c) >> 11;
  return ((double)x + 0.5) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------------ */
/* CBMC helpers shared by the MC driver and the batched Widom                */
/* ------------------------------------------------------------------------ */

/* Rotate_Quaternions (mc_utilities.h:423-457): uniform random rotation from
 * three uniforms (Shoemake / Kuffner). */
static void rotate_quaternion(double v[3], double u, double vv, double w)
{
  const double pi = 3.14159265358979323846;
  const double q0 = sqrt(1 - u) * sin(2 * pi * vv);
  const double q1 = sqrt(1 - u) * cos(2 * pi * vv);
  const double q2 = sqrt(u) * sin(2 * pi * w);
  const double q3 = sqrt(u) * cos(2 * pi * w);
  double rot[9];
  const double a01 = q0 * q1, a02 = q0 * q2, a03 = q0 * q3;
  const double a11 = q1 * q1, a12 = q1 * q2, a13 = q1 * q3;
  const double a22 = q2 * q2, a23 = q2 * q3, a33 = q3 * q3;
  rot[0] = 1.0 - 2.0 * (a22 + a33);
  rot[1] = 2.0 * (a12 - a03);
  rot[2] = 2.0 * (a13 + a02);
  rot[3] = 2.0 * (a12 + a03);
  rot[4] = 1.0 - 2.0 * (a11 + a33);
  rot[5] = 2.0 * (a23 - a01);
  rot[6] = 2.0 * (a13 - a02);
  rot[7] = 2.0 * (a23 + a01);
  rot[8] = 1.0 - 2.0 * (a11 + a22);
  const double r = v[0] * rot[0] + v[1] * rot[1] + v[2] * rot[2];
  const double s = v[0] * rot[3] + v[1] * rot[4] + v[2] * rot[5];
  const double c = v[0] * rot[6] + v[1] * rot[7] + v[2] * rot[8];
  v[0] = r; v[1] = s; v[2] = c;
}

/* Exported for the uniformity test of the orientation sampler. */
void or_rotate_quaternion(double v[3], double u, double vv, double w) { rotate_quaternion(v, u, vv, w); }

/* RotationAroundAxis (mc_utilities.h:459-483) */
static void rotate_axis(double p[3], double theta, double ax, double ay, double az)
{
  double c = cos(theta), w = 1.0 - c, s = sin(theta), rot[9];
  rot[0] = ax * ax * w + c;      rot[1] = ax * ay * w + az * s; rot[2] = ax * az * w - ay * s;
  rot[3] = ax * ay * w - az * s; rot[4] = ay * ay * w + c;      rot[5] = ay * az * w + ax * s;
  rot[6] = ax * az * w + ay * s; rot[7] = ay * az * w - ax * s; rot[8] = az * az * w + c;
  double x = p[0] * rot[0] + p[1] * rot[1] + p[2] * rot[2];
  double y = p[0] * rot[3] + p[1] * rot[4] + p[2] * rot[5];
  double z = p[0] * rot[6] + p[1] * rot[7] + p[2] * rot[8];
  p[0] = x; p[1] = y; p[2] = z;
}

/* SelectTrialPosition (mc_widom.h:14-39) with an externally drawn uniform. */
static int select_trial(const double* logw, int n, double u)
{
  double largest = logw[0];
  for (int i = 1; i < n; i++) if (logw[i] > largest) largest = logw[i];
  double shifted[1024];
  double sum = 0.0;
  for (int i = 0; i < n; i++) { shifted[i] = exp(logw[i] - largest); sum += shifted[i]; }
  int sel = 0;
  double cumw = shifted[0];
  double ws = u * sum;
  while (cumw < ws) cumw += shifted[++sel];
  return sel;
}

typedef struct {
  int    n_keep;           /* surviving (non-overlapping) trials          */
  int    keep[1024];       /* their trial indices                        */
  double rosen[1024];      /* -beta * (vdw [+ real]) of each survivor    */
  double e4[1024][4];      /* hg_vdw, hg_real, gg_vdw, gg_real           */
} trial_set;

/* Host_sum_Widom_HGGG_SEPARATE (mc_widom.h:41-87): per-trial sums over the
 * chain sites, survivors, log Boltzmann factors. */
static void sum_trials(const double* site4, const uint8_t* site_ov, int n_trials, int chain,
                       double beta, int bias, trial_set* ts)
{
  ts->n_keep = 0;
  for (int t = 0; t < n_trials; t++) {
    int ov = 0;
    double e[4] = {0, 0, 0, 0};
    for (int a = 0; a < chain; a++) {
      int s = t * chain + a;
      if (site_ov[s]) ov = 1;
      for (int k = 0; k < 4; k++) e[k] += site4[4 * s + k];
    }
    if (ov) continue;
    int i = ts->n_keep++;
    ts->keep[i] = t;
    for (int k = 0; k < 4; k++) ts->e4[i][k] = e[k];
    double tot = e[0] + e[2];
    if (bias) tot += e[1] + e[3];
    ts->rosen[i] = -beta * tot;
  }
}

static double sum_exp(const trial_set* ts)
{
  double s = 0.0;
  for (int i = 0; i < ts->n_keep; i++) s += exp(ts->rosen[i]);
  return s;
}

/* ------------------------------------------------------------------------ */
/* Flexible linear chains (no reference counterpart: the reference grows     */
/* rigid molecules only, read_data.cpp:2104-2116)                            */
/* ------------------------------------------------------------------------ */

/* Draws of bead i, trial j of a batched move: the block starting at
 * D_FLEX + ((i - 1) * 32 + j) * 1024.  Bead 1: draws 0, 1 (cos of the polar
 * angle, azimuth).  Bend angle (beads 2, 3): attempt a takes draws 3a, 3a + 1
 * (a Box-Muller normal) and 3a + 2 (acceptance), a < 160.  Dihedral (bead 3):
 * attempt a takes 512 + 2a (angle) and 513 + 2a (acceptance), a < 250.
 * Azimuth of bead 2: draw 1023.  A sampler that runs out of attempts returns
 * theta0 (bend) or pi (dihedral); for the butane model at 200-400 K that has
 * probability below 1e-30.  The uniform of step i's selection is draw
 * D_FLEX - 8 + i. */
enum { D_FLEX = 4096, FLEX_BLOCK = 1024, FLEX_BEND_TRIES = 160, FLEX_TORS_TRIES = 250 };

static double flex_u(uint64_t seed, uint64_t stream, uint64_t move, int i, int j, int d)
{
  return or_philox_uniform(seed, stream, move, (uint32_t)(D_FLEX + ((i - 1) * 32 + j) * FLEX_BLOCK + d));
}

/* Bend angle from sin(theta) exp(-beta k/2 (theta - theta0)^2) on (0, pi):
 * a normal proposal N(theta0, 1/sqrt(beta k)) carries the Boltzmann factor,
 * so accepting it with probability sin(theta) leaves the density above. */
static double flex_bend(const or_mc* mc, int comp, int i, uint64_t seed, uint64_t stream, uint64_t move, int j)
{
  const double pi = 3.14159265358979323846;
  const double th0 = mc->bend_theta0[comp][i];
  const double sd = 1.0 / sqrt(mc->beta * mc->bend_k[comp][i]);
  for (int a = 0; a < FLEX_BEND_TRIES; a++) {
    const double u1 = flex_u(seed, stream, move, i, j, 3 * a);
    const double u2 = flex_u(seed, stream, move, i, j, 3 * a + 1);
    const double ua = flex_u(seed, stream, move, i, j, 3 * a + 2);
    const double th = th0 + sd * sqrt(-2.0 * log(u1)) * cos(2.0 * pi * u2);
    if (th > 0.0 && th < pi && ua < sin(th)) return th;
  }
  return th0;
}

/* Dihedral from exp(-beta U(phi)) on [0, 2 pi): uniform proposals accepted
 * with exp(-beta (U - U_lb)), U_lb = t0 + min(0, 2 t1) + min(0, 2 t2) +
 * min(0, 2 t3) <= U everywhere. */
static double flex_dihedral(const or_mc* mc, int comp, int i, uint64_t seed, uint64_t stream, uint64_t move,
                            int j)
{
  const double pi = 3.14159265358979323846;
  const double* t = mc->torsion[comp][i];
  const double lb = t[0] + fmin(0.0, 2.0 * t[1]) + fmin(0.0, 2.0 * t[2]) + fmin(0.0, 2.0 * t[3]);
  for (int a = 0; a < FLEX_TORS_TRIES; a++) {
    const double phi = 2.0 * pi * flex_u(seed, stream, move, i, j, 512 + 2 * a);
    const double ua = flex_u(seed, stream, move, i, j, 513 + 2 * a);
    const double u = t[0] + t[1] * (1.0 + cos(phi)) + t[2] * (1.0 - cos(2.0 * phi)) + t[3] * (1.0 + cos(3.0 * phi));
    if (ua < exp(-mc->beta * (u - lb))) return phi;
  }
  return pi;
}

static void unit3(double v[3])
{
  const double r = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= r; v[1] /= r; v[2] /= r;
}

static void cross3(const double a[3], const double b[3], double out[3])
{
  out[0] = a[1] * b[2] - a[2] * b[1];
  out[1] = a[2] * b[0] - a[0] * b[2];
  out[2] = a[0] * b[1] - a[1] * b[0];
}

/* Bead i from beads i-1 (C), i-2 (B), i-3 (A) and (bond, bend, dihedral):
 * C + bond (-cos th bc + sin th (cos phi m + sin phi n)), bc = (C - B)/|C - B|,
 * n = unit(AB x bc), m = n x bc; phi is then the IUPAC dihedral A-B-C-D.
 * Bead 2 has no A: n is any unit vector normal to bc (built from the axis
 * least aligned with bc) and phi a uniform azimuth. */
void or_flex_trial(const or_mc* mc, int comp, int i, const double* prev, uint64_t seed, uint64_t stream,
                   uint64_t move, int j, double out[3])
{
  const double pi = 3.14159265358979323846;
  const double bl = mc->bond[comp][i];
  const double* C = prev + 3 * (i - 1);
  if (i == 1) {
    const double z = 2.0 * flex_u(seed, stream, move, i, j, 0) - 1.0;
    const double ph = 2.0 * pi * flex_u(seed, stream, move, i, j, 1);
    const double r = sqrt(fmax(0.0, 1.0 - z * z));
    out[0] = C[0] + bl * r * cos(ph);
    out[1] = C[1] + bl * r * sin(ph);
    out[2] = C[2] + bl * z;
    return;
  }
  const double* B = prev + 3 * (i - 2);
  double bc[3] = {C[0] - B[0], C[1] - B[1], C[2] - B[2]};
  unit3(bc);
  double n[3], m[3];
  if (i == 2) {
    double ax[3] = {0.0, 0.0, 0.0};
    const double fx = fabs(bc[0]), fy = fabs(bc[1]), fz = fabs(bc[2]);
    ax[(fx <= fy && fx <= fz) ? 0 : (fy <= fz ? 1 : 2)] = 1.0;
    cross3(bc, ax, n);
  } else {
    const double* A = prev + 3 * (i - 3);
    const double ab[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    cross3(ab, bc, n);
  }
  unit3(n);
  cross3(n, bc, m);
  const double th = flex_bend(mc, comp, i, seed, stream, move, j);
  const double ph = (i == 2) ? 2.0 * pi * flex_u(seed, stream, move, i, j, FLEX_BLOCK - 1)
                             : flex_dihedral(mc, comp, i, seed, stream, move, j);
  const double d0 = -bl * cos(th), d1 = bl * sin(th) * cos(ph), d2 = bl * sin(th) * sin(ph);
  for (int k = 0; k < 3; k++) out[k] = C[k] + d0 * bc[k] + d1 * m[k] + d2 * n[k];
}

/* ------------------------------------------------------------------------ */
/* Batched Widom insertions (counter-based streams)                          */
/* ------------------------------------------------------------------------ */

/* One CBMC Widom insertion, the semantics of Insertion_Body
 * (mc_swap_utilities.h:3-133) with Widom_Move_FirstBead_PARTIAL /
 * Widom_Move_Chain_PARTIAL (mc_widom.h:305-614); the trial positions and
 * orientations come from Philox draws instead of the reference's shared pool.
 * Draw layout for insertion i: 3k..3k+2 trial k position; 3K first-bead
 * selection; 3K+1+3o..3K+3+3o orientation o; 3K+3O+1 orientation selection. */
int or_widom_batch(const or_box* b, const or_atoms* comps, int ncomp, int comp,
                   const double* sk_ads, const double* sk_fw,
                   double beta, int K, int O, int bias,
                   double excl_intra, double excl_self, double tail_de,
                   uint64_t seed, uint64_t stream, int64_t first_index, int n_insert,
                   or_widom_out* out)
{
  const or_atoms* A = &comps[comp];
  int M = A->molsize;
  int chain = M - 1;
  int64_t new_molid = A->n_mol;
  if (K > 1024 || O * (chain > 0 ? chain : 1) > 1024) return -1;
  /* insertions are independent (each draws from its own counter-based
   * stream and writes its own record): OpenMP over insertions, per-thread
   * work buffers; results do not depend on the thread count */
#pragma omp parallel
  {
  double* sp = (double*)malloc(sizeof(double) * 3 * 1024);
  int* st = (int*)malloc(sizeof(int) * 1024);
  double* sq = (double*)malloc(sizeof(double) * 1024);
  double* s4 = (double*)malloc(sizeof(double) * 4 * 1024);
  uint8_t* sov = (uint8_t*)malloc(1024);
  trial_set* ts = (trial_set*)malloc(sizeof(trial_set));
#pragma omp for schedule(dynamic, 8)
  for (int n = 0; n < n_insert; n++) {
    uint64_t idx = (uint64_t)(first_index + n);
    or_widom_out* o = &out[n];
    memset(o, 0, sizeof(*o));
    /* first bead */
    for (int k = 0; k < K; k++) {
      double u0 = or_philox_uniform(seed, stream, idx, 3 * k);
      double u1 = or_philox_uniform(seed, stream, idx, 3 * k + 1);
      double u2 = or_philox_uniform(seed, stream, idx, 3 * k + 2);
      sp[3 * k] = b->cell[0] * u0; sp[3 * k + 1] = b->cell[4] * u1; sp[3 * k + 2] = b->cell[8] * u2;
      st[k] = A->type[0]; sq[k] = A->charge[0];
    }
    or_site_energies(b, comps, ncomp, K, sp, st, sq, comp, new_molid, -1, -1, s4, sov);
    sum_trials(s4, sov, K, 1, beta, bias, ts);
    if (ts->n_keep == 0) continue;
    int sel = select_trial(ts->rosen, ts->n_keep, or_philox_uniform(seed, stream, idx, 3 * K));
    double rb = sum_exp(ts);
    if (rb < 1e-150) continue;
    int fb = ts->keep[sel];
    double W = rb / (double)K;
    if (!bias) W *= exp(-beta * (ts->e4[sel][1] + ts->e4[sel][3]));
    double e_fb[4] = {ts->e4[sel][0], ts->e4[sel][1], ts->e4[sel][2], ts->e4[sel][3]};
    double fbpos[3] = {sp[3 * fb], sp[3 * fb + 1], sp[3 * fb + 2]};
    double mol_pos[3 * 64];
    double mol_q[64];
    mol_pos[0] = fbpos[0]; mol_pos[1] = fbpos[1]; mol_pos[2] = fbpos[2];
    mol_q[0] = A->charge[0];
    double e_ch[4] = {0, 0, 0, 0};
    int orient = 0;
    if (chain > 0) {
      for (int t = 0; t < O; t++) {
        double u = or_philox_uniform(seed, stream, idx, 3 * K + 1 + 3 * t);
        double v = or_philox_uniform(seed, stream, idx, 3 * K + 2 + 3 * t);
        double w = or_philox_uniform(seed, stream, idx, 3 * K + 3 + 3 * t);
        for (int a = 0; a < chain; a++) {
          double vec[3] = {A->pos[3 * (1 + a)] - A->pos[0], A->pos[3 * (1 + a) + 1] - A->pos[1],
                           A->pos[3 * (1 + a) + 2] - A->pos[2]};
          rotate_quaternion(vec, u, v, w);
          int s = t * chain + a;
          sp[3 * s] = fbpos[0] + vec[0]; sp[3 * s + 1] = fbpos[1] + vec[1]; sp[3 * s + 2] = fbpos[2] + vec[2];
          st[s] = A->type[1 + a]; sq[s] = A->charge[1 + a];
        }
      }
      or_site_energies(b, comps, ncomp, O * chain, sp, st, sq, comp, new_molid, -1, -1, s4, sov);
      sum_trials(s4, sov, O, chain, beta, bias, ts);
      if (ts->n_keep == 0) continue;
      int osel = select_trial(ts->rosen, ts->n_keep, or_philox_uniform(seed, stream, idx, 3 * K + 3 * O + 1));
      double rc = sum_exp(ts);
      if (rc < 1e-150) continue;
      orient = ts->keep[osel];
      double Wc = rc / (double)O;
      if (!bias) Wc *= exp(-beta * (ts->e4[osel][1] + ts->e4[osel][3]));
      W *= Wc;
      for (int k = 0; k < 4; k++) e_ch[k] = ts->e4[osel][k];
      for (int a = 0; a < chain; a++) {
        int s = orient * chain + a;
        mol_pos[3 * (1 + a)] = sp[3 * s]; mol_pos[3 * (1 + a) + 1] = sp[3 * s + 1]; mol_pos[3 * (1 + a) + 2] = sp[3 * s + 2];
        mol_q[1 + a] = A->charge[1 + a];
      }
    }
    if (W <= 1e-150) continue;
    double ew = 0.0;
    int has_q = 0;
    for (int a = 0; a < M; a++) if (fabs(A->charge[a]) > 0.0) has_q = 1;
    if (!b->no_charges && has_q) {
      double same, cross;
      or_ewald_delta(b, sk_ads, sk_fw, 0, NULL, NULL, M, mol_pos, mol_q, &same, &cross, NULL);
      same -= (excl_intra + excl_self);
      ew = same + 2.0 * cross;
      W *= exp(-beta * ew);
    }
    W *= exp(-beta * tail_de);
    o->rosen = W;
    o->vdw = e_fb[0] + e_fb[2] + e_ch[0] + e_ch[2];
    o->real = e_fb[1] + e_fb[3] + e_ch[1] + e_ch[3];
    o->ewald = ew;
    o->energy_total = o->vdw + o->real + ew + tail_de;
    o->first_bead = fb;
    o->orientation = orient;
    o->ok = 1;
  }
  free(sp); free(st); free(sq); free(s4); free(sov); free(ts);
  }
  return 0;
}

/* Threads of the OpenMP loops (or_widom_batch); n < 1 keeps the default. */
int or_set_threads(int n)
{
#ifdef _OPENMP
  if (n >= 1) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Monte Carlo driver in the reference's draw order                          */
/* ------------------------------------------------------------------------ */

/* RandomNumber (data_struct.h:1287-1346): a pool of double3 filled from
 * rand(); pool[0] is overwritten with {2.3, 4.5, 6.7} by Aaccess_device_random
 * at setup and stays so until the first refill. */
static double* g_pool = NULL;
static size_t g_pool_size = 0, g_pool_offset = 0;

static void pool_reset(void)
{
  g_pool_offset = 0;
  for (size_t i = 0; i < g_pool_size; i++) {
    g_pool[3 * i] = or_rand_uniform();
    g_pool[3 * i + 1] = or_rand_uniform();
    g_pool[3 * i + 2] = or_rand_uniform();
  }
  for (size_t i = g_pool_size * 3; i < 1000000; i++) (void)or_rand_uniform();
}
static void pool_check(size_t n) { if (g_pool_offset + n >= g_pool_size) pool_reset(); }

int or_mc_setup(or_mc* mc)
{
  free(g_pool);
  g_pool_size = (size_t)mc->pool_size;
  g_pool = (double*)malloc(sizeof(double) * 3 * g_pool_size);
  if (!g_pool) return -1;
  or_srand(mc->seed);
  pool_reset();
  g_pool[0] = 2.3; g_pool[1] = 4.5; g_pool[2] = 6.7;
  return 0;
}
void or_mc_free(void) { free(g_pool); g_pool = NULL; g_pool_size = 0; }

static double energy_total(const or_energy* e)
{
  return e->hh_vdw + e->hg_vdw + e->gg_vdw + e->hh_real + e->hg_real + e->gg_real + e->hh_ewald +
         e->hg_ewald + e->gg_ewald + e->tail;
}
static void energy_add(or_energy* a, const or_energy* b)
{
  a->hh_vdw += b->hh_vdw; a->hg_vdw += b->hg_vdw; a->gg_vdw += b->gg_vdw;
  a->hh_real += b->hh_real; a->hg_real += b->hg_real; a->gg_real += b->gg_real;
  a->hh_ewald += b->hh_ewald; a->hg_ewald += b->hg_ewald; a->gg_ewald += b->gg_ewald;
  a->tail += b->tail;
}

typedef struct {
  const or_box* b; or_atoms* comps; int ncomp; or_mc* mc; double* sk_ads; double* sk_fw; double* sk_tmp;
  /* batched form: the uniforms of move `move` come from the Philox stream
   * (seed, stream, move) by draw index instead of the shared pool */
  int philox; uint64_t seed, stream, move;
  /* creation phase (InsertionMove::CreateMolecule, move_struct.h:96-106):
   * insertions accepted with the uniform 1e-100, no counters */
  int create;
} mc_ctx;

static double mc_draw(const mc_ctx* x, int d) { return or_philox_uniform(x->seed, x->stream, x->move, (uint32_t)d); }

/* Draw layout of one move in the batched form (the engine's batch_mc.hip). */
enum { D_COMP = 0, D_MOL = 1, D_MOVE = 2, D_COIN = 3, D_ACC = 4, D_SEL_FB = 5, D_SEL_OR = 6, D_VEC = 8,
       D_ID_OLD = 11, D_ID_NEW = 12, D_ID_MOL = 13, D_TRIAL = 16 };

static int comp_has_charge(const mc_ctx* x, int c) { return x->mc->has_charge[c] && !x->b->no_charges; }

/* SingleBody_Prepare / _Calculation / _Acceptance (mc_single_particle.h:10-293),
 * get_new_position (mc_utilities.h:485-606). */
static void single_body(mc_ctx* x, int comp, int mol, int move, or_trace* tr)
{
  or_atoms* A = &x->comps[comp];
  or_mc* mc = x->mc;
  int M = A->molsize;
  int start = mol * M;
  if (move == OR_TRANSLATION) mc->trans_total[comp]++; else mc->rot_total[comp]++;
  double Rp[3];
  const double* R = Rp;
  if (x->philox) {
    for (int d = 0; d < 3; d++) Rp[d] = mc_draw(x, D_VEC + d);
  } else {
    pool_check((size_t)M);
    R = &g_pool[3 * g_pool_offset];
    g_pool_offset += (size_t)M;
  }
  double oldp[3 * 64], newp[3 * 64], q[64];
  int ty[64];
  for (int i = 0; i < M; i++) {
    for (int d = 0; d < 3; d++) oldp[3 * i + d] = A->pos[3 * (start + i) + d];
    q[i] = A->charge[start + i];
    ty[i] = A->type[start + i];
  }
  if (move == OR_TRANSLATION) {
    for (int i = 0; i < M; i++)
      for (int d = 0; d < 3; d++)
        newp[3 * i + d] = oldp[3 * i + d] + mc->max_translation[comp][d] * 2.0 * (R[d] - 0.5);
  } else {
    double ang[3];
    for (int d = 0; d < 3; d++) ang[d] = mc->max_rotation[comp][d] * 2.0 * (R[d] - 0.5);
    for (int i = 0; i < M; i++) {
      double p[3] = {oldp[3 * i] - oldp[0], oldp[3 * i + 1] - oldp[1], oldp[3 * i + 2] - oldp[2]};
      rotate_axis(p, ang[0], 1.0, 0.0, 0.0);
      rotate_axis(p, ang[1], 0.0, 1.0, 0.0);
      rotate_axis(p, ang[2], 0.0, 0.0, 1.0);
      for (int d = 0; d < 3; d++) newp[3 * i + d] = p[d] + oldp[d];
    }
  }
  int64_t molid = A->molid[start];
  double en[4 * 64], eo[4 * 64];
  uint8_t on[64], oo[64];
  or_site_energies(x->b, x->comps, x->ncomp, M, newp, ty, q, comp, molid, -1, -1, en, on);
  or_site_energies(x->b, x->comps, x->ncomp, M, oldp, ty, q, comp, molid, -1, -1, eo, oo);
  int overlap = 0;
  for (int i = 0; i < M; i++) if (on[i]) overlap = 1;
  or_energy de;
  memset(&de, 0, sizeof(de));
  double pacc = 0.0;
  if (!overlap) {
    for (int i = 0; i < M; i++) {
      de.hg_vdw += en[4 * i] - eo[4 * i];
      de.hg_real += en[4 * i + 1] - eo[4 * i + 1];
      de.gg_vdw += en[4 * i + 2] - eo[4 * i + 2];
      de.gg_real += en[4 * i + 3] - eo[4 * i + 3];
    }
    if (comp_has_charge(x, comp)) {
      double same, cross;
      or_ewald_delta(x->b, x->sk_ads, x->sk_fw, M, oldp, q, M, newp, q, &same, &cross, x->sk_tmp);
      de.gg_ewald = same;
      de.hg_ewald = 2.0 * cross;
    }
    pacc = 1.0 * exp(-mc->beta * energy_total(&de));
  }
  int acc = 0;
  if (!overlap) {
    double u = x->philox ? mc_draw(x, D_ACC) : or_rand_uniform();
    if (u < pacc) acc = 1;
  }
  if (acc) {
    for (int i = 0; i < M; i++)
      for (int d = 0; d < 3; d++) A->pos[3 * (start + i) + d] = newp[3 * i + d];
    if (comp_has_charge(x, comp)) memcpy(x->sk_ads, x->sk_tmp, sizeof(double) * 2 * or_nvec(x->b));
    if (move == OR_TRANSLATION) mc->trans_acc[comp]++; else mc->rot_acc[comp]++;
    energy_add(&mc->running, &de);
  } else {
    memset(&de, 0, sizeof(de));
  }
  tr->accepted = acc; tr->overlap = overlap; tr->pacc = pacc; tr->de = de;
}

/* CBMC insertion / deletion / Widom: Insertion_Body and Deletion_Body
 * (mc_swap_utilities.h:3-225), InsertionMove / DeletionMove (move_struct.h:4-168),
 * GetPrefactor (mc_utilities.h:315-351), AcceptInsertion / AcceptDeletion
 * (mc_utilities.h:357-417). */
static void cbmc_move(mc_ctx* x, int comp, int mol, int move, or_trace* tr)
{
  or_atoms* A = &x->comps[comp];
  or_mc* mc = x->mc;
  const or_box* b = x->b;
  int M = A->molsize, chain = M - 1;
  int K = mc->n_trial_pos, O = mc->n_trial_orient;
  int del = (move == OR_DELETION);
  int64_t N = A->n_mol;
  int64_t sel_molid = del ? (int64_t)mol : N;      /* TempVal.molecule for deletion */
  int start = del ? mol * M : 0;
  double beta = mc->beta;
  if (move == OR_INSERTION && !x->create) mc->ins_total[comp]++;
  if (move == OR_DELETION) mc->del_total[comp]++;
  if (move == OR_WIDOM) mc->widom_total[comp]++;

  static double sp[3 * 1024], sq[1024], s4[4 * 1024];
  static int st[1024];
  static uint8_t sov[1024];
  static trial_set ts;
  or_energy en;
  memset(&en, 0, sizeof(en));
  tr->accepted = 0; tr->overlap = 0; tr->pacc = 0.0; tr->rosen = 0.0;
  memset(&tr->de, 0, sizeof(tr->de));

  /* first bead: Widom_Move_FirstBead_PARTIAL (mc_widom.h:385-507),
   * get_random_trial_position (mc_widom.h:122-213) */
  if (!x->philox) pool_check((size_t)K);
  for (int k = 0; k < K; k++) {
    double Rp[3];
    const double* R = Rp;
    if (x->philox) for (int d = 0; d < 3; d++) Rp[d] = mc_draw(x, D_TRIAL + 3 * k + d);
    else R = &g_pool[3 * (g_pool_offset + k)];
    if (del && k == 0) {
      sp[0] = A->pos[3 * start]; sp[1] = A->pos[3 * start + 1]; sp[2] = A->pos[3 * start + 2];
    } else {
      sp[3 * k] = b->cell[0] * R[0]; sp[3 * k + 1] = b->cell[4] * R[1]; sp[3 * k + 2] = b->cell[8] * R[2];
    }
    st[k] = A->type[start]; sq[k] = A->charge[start];
  }
  if (!x->philox) g_pool_offset += (size_t)K;
  or_site_energies(b, x->comps, x->ncomp, K, sp, st, sq, comp, sel_molid, -1, -1, s4, sov);
  sum_trials(s4, sov, K, 1, beta, mc->vdw_real_bias, &ts);
  /* CBMC_FirstBead_Finish (mc_widom.h:305-383) */
  int sel = 0;
  double rb;
  if (!del) {
    if (ts.n_keep == 0) goto fail;
    sel = select_trial(ts.rosen, ts.n_keep, x->philox ? mc_draw(x, D_SEL_FB) : or_rand_uniform());
    rb = sum_exp(&ts);
    if (rb < 1e-150) goto fail;
  } else {
    sel = 0;
    rb = sum_exp(&ts);
  }
  double W = rb / (double)K;
  if (ts.n_keep > 0 && !mc->vdw_real_bias) W *= exp(-beta * (ts.e4[sel][1] + ts.e4[sel][3]));
  if (W <= 1e-150) goto fail;
  int fbt = ts.keep[sel];
  en.hg_vdw = ts.e4[sel][0]; en.hg_real = ts.e4[sel][1]; en.gg_vdw = ts.e4[sel][2]; en.gg_real = ts.e4[sel][3];
  double mol_pos[3 * 64], mol_q[64];
  int mol_t[64];
  mol_pos[0] = sp[3 * fbt]; mol_pos[1] = sp[3 * fbt + 1]; mol_pos[2] = sp[3 * fbt + 2];
  mol_q[0] = sq[fbt]; mol_t[0] = st[fbt];
  if (chain > 0 && mc->flexible[comp]) {
    /* flexible chain: bead i of O trial positions from the bonded
     * distributions given the beads already placed (deletion: trial 0 is the
     * molecule's own bead), selected on the Boltzmann factor of its external
     * energy; the step weight is that factor's trial average */
    for (int i = 1; i < M; i++) {
      for (int t = 0; t < O; t++) {
        if (del && t == 0) {
          for (int d = 0; d < 3; d++) sp[3 * t + d] = A->pos[3 * (start + i) + d];
        } else {
          or_flex_trial(mc, comp, i, mol_pos, x->seed, x->stream, x->move, t, &sp[3 * t]);
        }
        st[t] = A->type[start + i]; sq[t] = A->charge[start + i];
      }
      or_site_energies(b, x->comps, x->ncomp, O, sp, st, sq, comp, sel_molid, -1, -1, s4, sov);
      sum_trials(s4, sov, O, 1, beta, mc->vdw_real_bias, &ts);
      int osel = 0;
      double rc;
      if (ts.n_keep == 0) goto fail;
      if (!del) {
        osel = select_trial(ts.rosen, ts.n_keep, mc_draw(x, D_FLEX - 8 + i));
        rc = sum_exp(&ts);
        if (rc < 1e-150) goto fail;
      } else {
        rc = sum_exp(&ts);
      }
      double Wc = rc / (double)O;
      if (!mc->vdw_real_bias) Wc *= exp(-beta * (ts.e4[osel][1] + ts.e4[osel][3]));
      W *= Wc;
      if (W <= 1e-150) goto fail;
      int ot = ts.keep[osel];
      en.hg_vdw += ts.e4[osel][0]; en.hg_real += ts.e4[osel][1];
      en.gg_vdw += ts.e4[osel][2]; en.gg_real += ts.e4[osel][3];
      for (int d = 0; d < 3; d++) mol_pos[3 * i + d] = sp[3 * ot + d];
      mol_q[i] = sq[ot]; mol_t[i] = st[ot];
    }
  } else if (chain > 0) {
    /* Widom_Move_Chain_PARTIAL (mc_widom.h:509-614), get_random_trial_orientation
     * (mc_widom.h:215-303): template vectors from the component's first
     * molecule slot as it is now. */
    if (!x->philox) pool_check((size_t)O);
    for (int t = 0; t < O; t++) {
      double Rp[3];
      const double* R = Rp;
      if (x->philox) for (int d = 0; d < 3; d++) Rp[d] = mc_draw(x, D_TRIAL + 3 * K + 3 * t + d);
      else R = &g_pool[3 * (g_pool_offset + t)];
      for (int a = 0; a < chain; a++) {
        int s = t * chain + a;
        if (del && t == 0) {
          for (int d = 0; d < 3; d++) sp[3 * s + d] = A->pos[3 * (start + 1 + a) + d];
        } else {
          double vec[3] = {A->pos[3 * (1 + a)] - A->pos[0], A->pos[3 * (1 + a) + 1] - A->pos[1],
                           A->pos[3 * (1 + a) + 2] - A->pos[2]};
          rotate_quaternion(vec, R[0], R[1], R[2]);
          for (int d = 0; d < 3; d++) sp[3 * s + d] = mol_pos[d] + vec[d];
        }
        st[s] = A->type[start + 1 + a]; sq[s] = A->charge[start + 1 + a];
      }
    }
    if (!x->philox) g_pool_offset += (size_t)O;
    or_site_energies(b, x->comps, x->ncomp, O * chain, sp, st, sq, comp, sel_molid, -1, -1, s4, sov);
    sum_trials(s4, sov, O, chain, beta, mc->vdw_real_bias, &ts);
    int osel = 0;
    double rc;
    if (!del) {
      if (ts.n_keep == 0) goto fail;
      osel = select_trial(ts.rosen, ts.n_keep, x->philox ? mc_draw(x, D_SEL_OR) : or_rand_uniform());
      rc = sum_exp(&ts);
      if (rc < 1e-150) goto fail;
    } else {
      osel = 0;
      rc = sum_exp(&ts);
      if (ts.n_keep == 0) goto fail;  /* Rosenbluth 0 -> construction fails below */
    }
    double Wc = rc / (double)O;
    if (!mc->vdw_real_bias) Wc *= exp(-beta * (ts.e4[osel][1] + ts.e4[osel][3]));
    W *= Wc;
    if (W <= 1e-150) goto fail;
    int ot = ts.keep[osel];
    en.hg_vdw += ts.e4[osel][0]; en.hg_real += ts.e4[osel][1];
    en.gg_vdw += ts.e4[osel][2]; en.gg_real += ts.e4[osel][3];
    for (int a = 0; a < chain; a++) {
      int s = ot * chain + a;
      for (int d = 0; d < 3; d++) mol_pos[3 * (1 + a) + d] = sp[3 * s + d];
      mol_q[1 + a] = sq[s]; mol_t[1 + a] = st[s];
    }
  }
  double V = b->volume;
  double Nd = (double)N;
  double pref = del ? Nd / (beta * mc->mol_fraction[comp] * mc->pressure * mc->fugacity_coeff[comp] * V)
                    : beta * mc->mol_fraction[comp] * mc->pressure * mc->fugacity_coeff[comp] * V / (1.0 + Nd);
  if (comp_has_charge(x, comp)) {
    double same, cross;
    double excl = mc->excl_intra[comp] + mc->excl_self[comp];
    if (!del) {
      or_ewald_delta(b, x->sk_ads, x->sk_fw, 0, NULL, NULL, M, mol_pos, mol_q, &same, &cross, x->sk_tmp);
      same -= excl * 1.0;
      en.gg_ewald = same; en.hg_ewald = 2.0 * cross;
      W *= exp(-beta * (same + 2.0 * cross));
    } else {
      double op[3 * 64], oq[64];
      for (int i = 0; i < M; i++) {
        for (int d = 0; d < 3; d++) op[3 * i + d] = A->pos[3 * (start + i) + d];
        oq[i] = A->charge[start + i];
      }
      or_ewald_delta(b, x->sk_ads, x->sk_fw, M, op, oq, 0, NULL, NULL, &same, &cross, x->sk_tmp);
      same -= excl * (0.0 - 1.0);
      W /= exp(-beta * (same + 2.0 * cross));
      en.gg_ewald = -1.0 * same; en.hg_ewald = -1.0 * (2.0 * cross);
    }
  }
  if (mc->has_tail) {
    double te = or_tail_delta(b, mc->n_pseudo, mc->pseudo_per_mol[comp], del ? -1 : 1);
    if (!del) { W *= exp(-beta * te); en.tail = te; }
    else { W /= exp(-beta * te); en.tail = -te; }
  }
  tr->rosen = W;
  if (move == OR_WIDOM) {
    if (mc->production) { mc->widom_rosen_sum[comp] += W; mc->widom_rosen_n[comp] += 1.0; }
    tr->de = en;
    tr->pacc = 0.0;
    return;
  }
  double pacc = del ? pref * mc->ideal_rosen[comp] / W : pref * W / mc->ideal_rosen[comp];
  if (!del && W <= 1e-150) pacc = 0.0;
  tr->pacc = pacc;
  double u = x->create ? 1e-100 : (x->philox ? mc_draw(x, D_ACC) : or_rand_uniform());
  if (!(u < pacc)) return;  /* rejected: energy zero */
  /* accepted */
  tr->accepted = 1;
  int nv = or_nvec(b);
  if (!del) {
    int loc = (int)(N * M);
    if (A->n_mol + 1 > A->capacity) return;
    for (int i = 0; i < M; i++) {
      for (int d = 0; d < 3; d++) A->pos[3 * (loc + i) + d] = mol_pos[3 * i + d];
      A->charge[loc + i] = mol_q[i];
      A->type[loc + i] = mol_t[i];
      A->molid[loc + i] = N;
    }
    A->n_mol++;
    mc->total_molecules++;
    for (int t = 0; t < b->ntypes && t < OR_MAX_PSEUDO; t++) mc->n_pseudo[t] += mc->pseudo_per_mol[comp][t];
    if (!x->create) mc->ins_acc[comp]++;
  } else {
    int last = (int)((N - 1) * M);
    if (start != last) {
      for (int i = 0; i < M; i++) {
        for (int d = 0; d < 3; d++) A->pos[3 * (start + i) + d] = A->pos[3 * (last + i) + d];
        A->charge[start + i] = A->charge[last + i];
        A->type[start + i] = A->type[last + i];
      }
    }
    A->n_mol--;
    mc->total_molecules--;
    for (int t = 0; t < b->ntypes && t < OR_MAX_PSEUDO; t++) mc->n_pseudo[t] -= mc->pseudo_per_mol[comp][t];
    mc->del_acc[comp]++;
    /* take_negative */
    en.hg_vdw = -en.hg_vdw; en.hg_real = -en.hg_real; en.gg_vdw = -en.gg_vdw; en.gg_real = -en.gg_real;
    en.gg_ewald = -en.gg_ewald; en.hg_ewald = -en.hg_ewald; en.tail = -en.tail;
  }
  if (comp_has_charge(x, comp)) memcpy(x->sk_ads, x->sk_tmp, sizeof(double) * 2 * nv);
  tr->de = en;
  energy_add(&mc->running, &en);
  return;
fail:
  tr->rosen = 0.0;
  if (move == OR_WIDOM && mc->production) { mc->widom_rosen_n[comp] += 1.0; }
}

/* CBMC reinsertion: ReinsertionMove::Run (move_struct.h:170-407).  A new
 * configuration of molecule `mol` is grown as an insertion (K random first
 * beads, O orientations of the template about the selected bead, energies
 * excluding the molecule itself); the old one is retraced with one first-bead
 * trial (its own position, whose weight is added to the new growth's other
 * first-bead weights, StoredR, CBMC_FirstBead_Finish mc_widom.h:365-366) and
 * O orientations (trial 0 = its own).  W_new is multiplied by the Ewald
 * change of moving the molecule; Pacc = W_new / W_old.  Draws in the pool
 * order: K triples, selection, O triples, selection, 1 triple (retrace first
 * bead, unused), O triples (retrace), Metropolis uniform. */
static void reinsertion_move(mc_ctx* x, int comp, int mol, or_trace* tr)
{
  or_atoms* A = &x->comps[comp];
  or_mc* mc = x->mc;
  const or_box* b = x->b;
  const int M = A->molsize, chain = M - 1;
  const int K = mc->n_trial_pos, O = mc->n_trial_orient;
  const int start = mol * M;
  const int64_t molid = A->molid[start];
  const double beta = mc->beta;
  mc->reins_total[comp]++;
  static double sp[3 * 1024], sq[1024], s4[4 * 1024];
  static int st[1024];
  static uint8_t sov[1024];
  static trial_set ts;
  or_energy en_new, en_old, de;
  memset(&en_new, 0, sizeof(en_new));
  memset(&en_old, 0, sizeof(en_old));
  memset(&de, 0, sizeof(de));
  tr->accepted = 0; tr->overlap = 0; tr->pacc = 0.0; tr->rosen = 0.0;
  memset(&tr->de, 0, sizeof(tr->de));

  /* ---- growth of the new configuration (REINSERTION_INSERTION) ---- */
  if (!x->philox) pool_check((size_t)K);
  for (int k = 0; k < K; k++) {
    double Rp[3];
    const double* R = Rp;
    if (x->philox) for (int d = 0; d < 3; d++) Rp[d] = mc_draw(x, D_TRIAL + 3 * k + d);
    else R = &g_pool[3 * (g_pool_offset + k)];
    sp[3 * k] = b->cell[0] * R[0]; sp[3 * k + 1] = b->cell[4] * R[1]; sp[3 * k + 2] = b->cell[8] * R[2];
    st[k] = A->type[start]; sq[k] = A->charge[start];
  }
  if (!x->philox) g_pool_offset += (size_t)K;
  or_site_energies(b, x->comps, x->ncomp, K, sp, st, sq, comp, molid, -1, -1, s4, sov);
  sum_trials(s4, sov, K, 1, beta, mc->vdw_real_bias, &ts);
  if (ts.n_keep == 0) return;
  int sel = select_trial(ts.rosen, ts.n_keep, x->philox ? mc_draw(x, D_SEL_FB) : or_rand_uniform());
  double rb = sum_exp(&ts);
  if (rb < 1e-150) return;
  const double stored_r = rb - exp(ts.rosen[sel]);
  double Wn = rb / (double)K;
  if (!mc->vdw_real_bias) Wn *= exp(-beta * (ts.e4[sel][1] + ts.e4[sel][3]));
  if (Wn <= 1e-150) return;
  const int fbt = ts.keep[sel];
  en_new.hg_vdw = ts.e4[sel][0]; en_new.hg_real = ts.e4[sel][1];
  en_new.gg_vdw = ts.e4[sel][2]; en_new.gg_real = ts.e4[sel][3];
  double new_pos[3 * 64], old_pos[3 * 64], q[64];
  for (int i = 0; i < M; i++) {
    for (int d = 0; d < 3; d++) old_pos[3 * i + d] = A->pos[3 * (start + i) + d];
    q[i] = A->charge[start + i];
  }
  for (int d = 0; d < 3; d++) new_pos[d] = sp[3 * fbt + d];
  if (chain > 0) {
    if (!x->philox) pool_check((size_t)O);
    for (int t = 0; t < O; t++) {
      double Rp[3];
      const double* R = Rp;
      if (x->philox) for (int d = 0; d < 3; d++) Rp[d] = mc_draw(x, D_TRIAL + 3 * K + 3 * t + d);
      else R = &g_pool[3 * (g_pool_offset + t)];
      for (int a = 0; a < chain; a++) {
        const int s = t * chain + a;
        double vec[3] = {A->pos[3 * (1 + a)] - A->pos[0], A->pos[3 * (1 + a) + 1] - A->pos[1],
                         A->pos[3 * (1 + a) + 2] - A->pos[2]};
        rotate_quaternion(vec, R[0], R[1], R[2]);
        for (int d = 0; d < 3; d++) sp[3 * s + d] = new_pos[d] + vec[d];
        st[s] = A->type[start + 1 + a]; sq[s] = A->charge[start + 1 + a];
      }
    }
    if (!x->philox) g_pool_offset += (size_t)O;
    or_site_energies(b, x->comps, x->ncomp, O * chain, sp, st, sq, comp, molid, -1, -1, s4, sov);
    sum_trials(s4, sov, O, chain, beta, mc->vdw_real_bias, &ts);
    if (ts.n_keep == 0) return;
    const int osel = select_trial(ts.rosen, ts.n_keep, x->philox ? mc_draw(x, D_SEL_OR) : or_rand_uniform());
    const double rc = sum_exp(&ts);
    if (rc < 1e-150) return;
    double Wc = rc / (double)O;
    if (!mc->vdw_real_bias) Wc *= exp(-beta * (ts.e4[osel][1] + ts.e4[osel][3]));
    Wn *= Wc;
    if (Wn <= 1e-150) return;
    const int ot = ts.keep[osel];
    en_new.hg_vdw += ts.e4[osel][0]; en_new.hg_real += ts.e4[osel][1];
    en_new.gg_vdw += ts.e4[osel][2]; en_new.gg_real += ts.e4[osel][3];
    for (int a = 0; a < chain; a++)
      for (int d = 0; d < 3; d++) new_pos[3 * (1 + a) + d] = sp[3 * (ot * chain + a) + d];
  }

  /* ---- retrace of the old configuration (REINSERTION_RETRACE) ---- */
  if (!x->philox) { pool_check(1); g_pool_offset += 1; }      /* the retrace first bead takes one triple */
  for (int d = 0; d < 3; d++) sp[d] = old_pos[d];
  st[0] = A->type[start]; sq[0] = A->charge[start];
  or_site_energies(b, x->comps, x->ncomp, 1, sp, st, sq, comp, molid, -1, -1, s4, sov);
  sum_trials(s4, sov, 1, 1, beta, mc->vdw_real_bias, &ts);
  double Wo = (sum_exp(&ts) + stored_r) / (double)K;
  if (ts.n_keep > 0) {
    if (!mc->vdw_real_bias) Wo *= exp(-beta * (ts.e4[0][1] + ts.e4[0][3]));
    en_old.hg_vdw = ts.e4[0][0]; en_old.hg_real = ts.e4[0][1];
    en_old.gg_vdw = ts.e4[0][2]; en_old.gg_real = ts.e4[0][3];
  }
  if (chain > 0) {
    if (!x->philox) pool_check((size_t)O);
    for (int t = 0; t < O; t++) {
      double Rp[3];
      const double* R = Rp;
      if (x->philox) for (int d = 0; d < 3; d++) Rp[d] = mc_draw(x, D_TRIAL + 3 * K + 3 * O + 3 * t + d);
      else R = &g_pool[3 * (g_pool_offset + t)];
      for (int a = 0; a < chain; a++) {
        const int s = t * chain + a;
        if (t == 0) {
          for (int d = 0; d < 3; d++) sp[3 * s + d] = old_pos[3 * (1 + a) + d];
        } else {
          double vec[3] = {A->pos[3 * (1 + a)] - A->pos[0], A->pos[3 * (1 + a) + 1] - A->pos[1],
                           A->pos[3 * (1 + a) + 2] - A->pos[2]};
          rotate_quaternion(vec, R[0], R[1], R[2]);
          for (int d = 0; d < 3; d++) sp[3 * s + d] = old_pos[d] + vec[d];
        }
        st[s] = A->type[start + 1 + a]; sq[s] = A->charge[start + 1 + a];
      }
    }
    if (!x->philox) g_pool_offset += (size_t)O;
    or_site_energies(b, x->comps, x->ncomp, O * chain, sp, st, sq, comp, molid, -1, -1, s4, sov);
    sum_trials(s4, sov, O, chain, beta, mc->vdw_real_bias, &ts);
    double Wc = sum_exp(&ts) / (double)O;
    if (ts.n_keep > 0) {           /* SelectedTrial = 0 of the survivors (mc_widom.h:588-604) */
      if (!mc->vdw_real_bias) Wc *= exp(-beta * (ts.e4[0][1] + ts.e4[0][3]));
      en_old.hg_vdw += ts.e4[0][0]; en_old.hg_real += ts.e4[0][1];
      en_old.gg_vdw += ts.e4[0][2]; en_old.gg_real += ts.e4[0][3];
    }
    Wo *= Wc;
  }
  de.hg_vdw = en_new.hg_vdw - en_old.hg_vdw; de.hg_real = en_new.hg_real - en_old.hg_real;
  de.gg_vdw = en_new.gg_vdw - en_old.gg_vdw; de.gg_real = en_new.gg_real - en_old.gg_real;

  /* ---- Ewald change of moving the molecule (Calculate_AdjustRosenbluth) ---- */
  if (comp_has_charge(x, comp)) {
    double same, cross;
    or_ewald_delta(b, x->sk_ads, x->sk_fw, M, old_pos, q, M, new_pos, q, &same, &cross, x->sk_tmp);
    de.gg_ewald = same;
    de.hg_ewald = 2.0 * cross;
    Wn *= exp(-beta * (same + 2.0 * cross));
  }
  /* ---- acceptance (ReinsertionMove::Acceptance) ---- */
  const double u = x->philox ? mc_draw(x, D_ACC) : or_rand_uniform();
  const double pacc = Wn / Wo;
  tr->rosen = Wn;
  tr->pacc = pacc;
  if (Wn <= 1e-150 || !(u < pacc)) return;
  tr->accepted = 1;
  for (int i = 0; i < M; i++)
    for (int d = 0; d < 3; d++) A->pos[3 * (start + i) + d] = new_pos[3 * i + d];
  if (comp_has_charge(x, comp)) memcpy(x->sk_ads, x->sk_tmp, sizeof(double) * 2 * or_nvec(b));
  mc->reins_acc[comp]++;
  tr->de = de;
  energy_add(&mc->running, &de);
}

/* Identity swap: IdentitySwapMove (mc_swap_moves.h:199-426).  A molecule of
 * the OLD species is replaced by one of the NEW species grown on its first
 * bead: the NEW first bead sits at the OLD molecule's first bead (one trial,
 * copy_firstbead_to_new :178-181, its weight not divided by K), the NEW chain
 * takes O orientations of the NEW template (slot 0) about it, energies
 * without the OLD molecule (ExcludeList); the OLD molecule is retraced with
 * its own first bead (one trial) and O orientations (trial 0 its own).
 * W_new carries the Ewald change of removing the OLD sites and adding the
 * NEW ones (GPU_EwaldDifference_IdentitySwap, Ewald_Energy_Functions.h:582-632)
 * and the tail change with d = pseudo(NEW) - pseudo(OLD)
 * (TailCorrectionIdentitySwap, TailCorrection_Energy_Functions.h:85-108);
 * Pacc = pref(NEW, insertion) pref(OLD, deletion) (W_new / W_id,NEW) /
 * (W_old / W_id,OLD) with GetPrefactor (mc_utilities.h:315-345).  On
 * acceptance the OLD molecule is deleted (the last one into its hole) and the
 * NEW one appended, or, for the same species, moved in place.
 * Draws in the pool order: the species pair (two uniforms per try until the
 * OLD species holds a molecule), the OLD molecule, one triple (NEW first
 * bead, unused), O triples and a selection (NEW chain), one triple (OLD first
 * bead, unused), O triples (OLD chain), Metropolis. */
static void identity_swap_move(mc_ctx* x, or_trace* tr)
{
  or_mc* mc = x->mc;
  const or_box* b = x->b;
  or_atoms* comps = x->comps;
  const int ncomp = x->ncomp;
  const int K = mc->n_trial_pos, O = mc->n_trial_orient;
  const double beta = mc->beta;
  static double sp[3 * 1024], sq[1024], s4[4 * 1024];
  static int st[1024];
  static uint8_t sov[1024];
  static trial_set ts;
  tr->accepted = 0; tr->overlap = 0; tr->pacc = 0.0; tr->rosen = 0.0;
  memset(&tr->de, 0, sizeof(tr->de));
  if (mc->total_molecules - 1 == 0) return;          /* no adsorbates: no draws */
  int oc = 0, nc = 0;
  int64_t nold = 0;
  if (x->philox) {
    int held[8], nh = 0;
    for (int c = 1; c < ncomp && c < 8; c++)
      if (comps[c].n_mol > 0) held[nh++] = c;
    int i = (int)(mc_draw(x, D_ID_OLD) * (double)nh);
    if (i >= nh) i = nh - 1;
    oc = held[i];
    nc = 1 + (int)(mc_draw(x, D_ID_NEW) * (double)(ncomp - 1));
    if (nc >= ncomp) nc = ncomp - 1;
  } else {
    while (oc == 0 || oc >= ncomp || nc == 0 || nc >= ncomp || nold == 0) {
      oc = (int)(or_rand_uniform() * (double)(ncomp - 1)) + 1;
      nc = (int)(or_rand_uniform() * (double)(ncomp - 1)) + 1;
      nold = oc < ncomp ? comps[oc].n_mol : 0;
    }
  }
  or_atoms* AO = &comps[oc];
  or_atoms* AN = &comps[nc];
  const int oldmol = (int)((x->philox ? mc_draw(x, D_ID_MOL) : or_rand_uniform()) * (double)AO->n_mol);
  const int Mo = AO->molsize, Mn = AN->molsize;
  const int so = oldmol * Mo;
  const int64_t nnew = AN->n_mol;
  tr->comp = oc; tr->mol = oldmol;
  mc->idsw_total[oc][nc]++;
  or_energy de;
  memset(&de, 0, sizeof(de));
  double en_new[4] = {0, 0, 0, 0}, en_old[4] = {0, 0, 0, 0};
  double new_pos[3 * 64], new_q[64], old_pos[3 * 64], old_q[64];
  int new_t[64];
  for (int i = 0; i < Mo; i++) {
    for (int d = 0; d < 3; d++) old_pos[3 * i + d] = AO->pos[3 * (so + i) + d];
    old_q[i] = AO->charge[so + i];
  }
  for (int i = 0; i < Mn; i++) { new_q[i] = AN->charge[i]; new_t[i] = AN->type[i]; }

  /* ---- NEW first bead at the OLD molecule's first bead ---- */
  if (!x->philox) { pool_check(1); g_pool_offset += 1; }
  for (int d = 0; d < 3; d++) sp[d] = old_pos[d];
  st[0] = new_t[0]; sq[0] = new_q[0];
  or_site_energies(b, comps, ncomp, 1, sp, st, sq, nc, nnew, oc, oldmol, s4, sov);
  sum_trials(s4, sov, 1, 1, beta, mc->vdw_real_bias, &ts);
  if (ts.n_keep == 0) return;
  double Wn = sum_exp(&ts);
  if (Wn < 1e-150) return;
  if (!mc->vdw_real_bias) Wn *= exp(-beta * (ts.e4[0][1] + ts.e4[0][3]));
  if (Wn <= 1e-150) return;
  for (int k = 0; k < 4; k++) en_new[k] = ts.e4[0][k];
  for (int d = 0; d < 3; d++) new_pos[d] = old_pos[d];

  /* ---- NEW chain: O orientations of the NEW template ---- */
  if (Mn > 1) {
    const int chain = Mn - 1;
    if (!x->philox) pool_check((size_t)O);
    for (int t = 0; t < O; t++) {
      double Rp[3];
      const double* R = Rp;
      if (x->philox) for (int d = 0; d < 3; d++) Rp[d] = mc_draw(x, D_TRIAL + 3 * K + 3 * t + d);
      else R = &g_pool[3 * (g_pool_offset + t)];
      for (int a = 0; a < chain; a++) {
        const int s = t * chain + a;
        double vec[3] = {AN->pos[3 * (1 + a)] - AN->pos[0], AN->pos[3 * (1 + a) + 1] - AN->pos[1],
                         AN->pos[3 * (1 + a) + 2] - AN->pos[2]};
        rotate_quaternion(vec, R[0], R[1], R[2]);
        for (int d = 0; d < 3; d++) sp[3 * s + d] = new_pos[d] + vec[d];
        st[s] = new_t[1 + a]; sq[s] = new_q[1 + a];
      }
    }
    if (!x->philox) g_pool_offset += (size_t)O;
    or_site_energies(b, comps, ncomp, O * chain, sp, st, sq, nc, nnew, oc, oldmol, s4, sov);
    sum_trials(s4, sov, O, chain, beta, mc->vdw_real_bias, &ts);
    if (ts.n_keep == 0) return;
    const int osel = select_trial(ts.rosen, ts.n_keep, x->philox ? mc_draw(x, D_SEL_OR) : or_rand_uniform());
    const double rc = sum_exp(&ts);
    if (rc < 1e-150) return;
    double Wc = rc / (double)O;
    if (!mc->vdw_real_bias) Wc *= exp(-beta * (ts.e4[osel][1] + ts.e4[osel][3]));
    Wn *= Wc;
    if (Wn <= 1e-150) return;
    for (int k = 0; k < 4; k++) en_new[k] += ts.e4[osel][k];
    const int ot = ts.keep[osel];
    for (int a = 0; a < chain; a++)
      for (int d = 0; d < 3; d++) new_pos[3 * (1 + a) + d] = sp[3 * (ot * chain + a) + d];
  }

  /* ---- retrace of the OLD molecule: its own first bead ---- */
  if (!x->philox) { pool_check(1); g_pool_offset += 1; }
  for (int d = 0; d < 3; d++) sp[d] = old_pos[d];
  st[0] = AO->type[so]; sq[0] = old_q[0];
  or_site_energies(b, comps, ncomp, 1, sp, st, sq, oc, oldmol, -1, -1, s4, sov);
  sum_trials(s4, sov, 1, 1, beta, mc->vdw_real_bias, &ts);
  double Wo = sum_exp(&ts);
  if (ts.n_keep > 0) {
    if (!mc->vdw_real_bias) Wo *= exp(-beta * (ts.e4[0][1] + ts.e4[0][3]));
    for (int k = 0; k < 4; k++) en_old[k] = ts.e4[0][k];
  }
  /* ---- OLD chain: trial 0 its own orientation ---- */
  if (Mo > 1) {
    const int chain = Mo - 1;
    if (!x->philox) pool_check((size_t)O);
    for (int t = 0; t < O; t++) {
      double Rp[3];
      const double* R = Rp;
      if (x->philox) for (int d = 0; d < 3; d++) Rp[d] = mc_draw(x, D_TRIAL + 3 * K + 3 * O + 3 * t + d);
      else R = &g_pool[3 * (g_pool_offset + t)];
      for (int a = 0; a < chain; a++) {
        const int s = t * chain + a;
        if (t == 0) {
          for (int d = 0; d < 3; d++) sp[3 * s + d] = old_pos[3 * (1 + a) + d];
        } else {
          double vec[3] = {AO->pos[3 * (1 + a)] - AO->pos[0], AO->pos[3 * (1 + a) + 1] - AO->pos[1],
                           AO->pos[3 * (1 + a) + 2] - AO->pos[2]};
          rotate_quaternion(vec, R[0], R[1], R[2]);
          for (int d = 0; d < 3; d++) sp[3 * s + d] = old_pos[d] + vec[d];
        }
        st[s] = AO->type[so + 1 + a]; sq[s] = old_q[1 + a];
      }
    }
    if (!x->philox) g_pool_offset += (size_t)O;
    or_site_energies(b, comps, ncomp, O * chain, sp, st, sq, oc, oldmol, -1, -1, s4, sov);
    sum_trials(s4, sov, O, chain, beta, mc->vdw_real_bias, &ts);
    double Wc = sum_exp(&ts) / (double)O;
    if (ts.n_keep > 0) {
      if (!mc->vdw_real_bias) Wc *= exp(-beta * (ts.e4[0][1] + ts.e4[0][3]));
      for (int k = 0; k < 4; k++) en_old[k] += ts.e4[0][k];
    }
    Wo *= Wc;
  }
  de.hg_vdw = en_new[0] - en_old[0]; de.hg_real = en_new[1] - en_old[1];
  de.gg_vdw = en_new[2] - en_old[2]; de.gg_real = en_new[3] - en_old[3];

  /* ---- Ewald: OLD sites out, NEW sites in (charged species only) ---- */
  const int qo = comp_has_charge(x, oc), qn = comp_has_charge(x, nc);
  const int ewald = !b->no_charges && (qo || qn);
  if (ewald) {
    double same, cross;
    or_ewald_delta(b, x->sk_ads, x->sk_fw, qo ? Mo : 0, old_pos, old_q, qn ? Mn : 0, new_pos, new_q, &same, &cross,
                   x->sk_tmp);
    if (qn) same -= (mc->excl_intra[nc] + mc->excl_self[nc]) * 1.0;
    if (qo) same -= (mc->excl_intra[oc] + mc->excl_self[oc]) * -1.0;
    de.gg_ewald = same;
    de.hg_ewald = 2.0 * cross;
    Wn *= exp(-beta * (same + 2.0 * cross));
  }
  /* ---- tail ---- */
  if (mc->has_tail) {
    int64_t d[OR_MAX_PSEUDO];
    for (int t = 0; t < OR_MAX_PSEUDO; t++) d[t] = mc->pseudo_per_mol[nc][t] - mc->pseudo_per_mol[oc][t];
    de.tail = or_tail_delta(b, mc->n_pseudo, d, 1);
  }
  Wn *= exp(-beta * de.tail);

  /* ---- acceptance ---- */
  const double V = b->volume;
  double pref = beta * mc->mol_fraction[nc] * mc->pressure * mc->fugacity_coeff[nc] * V / (1.0 + (double)AN->n_mol);
  pref *= (double)AO->n_mol / (beta * mc->mol_fraction[oc] * mc->pressure * mc->fugacity_coeff[oc] * V);
  const double pacc = pref * (Wn / mc->ideal_rosen[nc]) / (Wo / mc->ideal_rosen[oc]);
  tr->rosen = Wn;
  tr->pacc = pacc;
  const double u = x->philox ? mc_draw(x, D_ACC) : or_rand_uniform();
  if (!(u < pacc)) return;
  if (nc != oc && AN->n_mol + 1 > AN->capacity) return;
  tr->accepted = 1;
  mc->idsw_acc[oc][nc]++;
  if (nc != oc) {
    const int last = (AO->n_mol - 1) * Mo;
    if (so != last) {
      for (int i = 0; i < Mo; i++) {
        for (int d = 0; d < 3; d++) AO->pos[3 * (so + i) + d] = AO->pos[3 * (last + i) + d];
        AO->charge[so + i] = AO->charge[last + i];
        AO->type[so + i] = AO->type[last + i];
      }
    }
    AO->n_mol--;
    const int loc = (int)(nnew * Mn);
    for (int i = 0; i < Mn; i++) {
      for (int d = 0; d < 3; d++) AN->pos[3 * (loc + i) + d] = new_pos[3 * i + d];
      AN->charge[loc + i] = new_q[i];
      AN->type[loc + i] = new_t[i];
      AN->molid[loc + i] = nnew;
    }
    AN->n_mol++;
    for (int t = 0; t < b->ntypes && t < OR_MAX_PSEUDO; t++) mc->n_pseudo[t] += mc->pseudo_per_mol[nc][t] - mc->pseudo_per_mol[oc][t];
  } else {
    for (int i = 0; i < Mo; i++)
      for (int d = 0; d < 3; d++) AO->pos[3 * (so + i) + d] = new_pos[3 * i + d];
  }
  if (ewald) memcpy(x->sk_ads, x->sk_tmp, sizeof(double) * 2 * or_nvec(b));
  tr->de = de;
  energy_add(&mc->running, &de);
}

/* Update_Max_Translation / Update_Max_Rotation (mc_utilities.h:612-665) */
static void update_max(double v[3], int64_t* acc, int64_t* tot, double hi)
{
  double ratio = (double)(*acc) / (double)(*tot);
  for (int d = 0; d < 3; d++) {
    v[d] *= (ratio > 0.5) ? 1.05 : 0.95;
    if (v[d] < 0.01) v[d] = 0.01;
    if (v[d] > hi) v[d] = hi;
  }
  *acc = 0; *tot = 0;
}

int or_mc_cycles(or_mc* mc, const or_box* b, or_atoms* comps, int ncomp,
                 double* sk_ads, double* sk_fw, int first_cycle, int n_cycles,
                 or_trace* trace, int max_trace)
{
  if (!g_pool || b->ntypes > OR_MAX_PSEUDO) return -1;
  for (int c = 1; c < ncomp && c < 8; c++)
    if (mc->flexible[c]) return -1;   /* flexible chains: batched form only */
  mc_ctx x;
  memset(&x, 0, sizeof(x));
  x.b = b; x.comps = comps; x.ncomp = ncomp; x.mc = mc; x.sk_ads = sk_ads; x.sk_fw = sk_fw;
  x.sk_tmp = (double*)calloc((size_t)2 * or_nvec(b), sizeof(double));
  int nt = 0;
  for (int cyc = first_cycle; cyc < first_cycle + n_cycles; cyc++) {
    /* Determine_Number_Of_Steps (axpy.cu:580-591) */
    int64_t steps = 20;
    if (steps < mc->total_molecules) steps = mc->total_molecules;
    if (mc->use_max_step && steps > mc->max_step_per_cycle) steps = mc->max_step_per_cycle;
    for (int64_t s = 0; s < steps; s++) {
      /* Select_Box_Component_Molecule (axpy.cu:74-101) */
      int comp = 0;
      while (!(comp >= 1 && mc->p_swap[comp] > 1e-10)) comp = (int)(or_rand_uniform() * (double)ncomp);
      int mol = (int)(or_rand_uniform() * (double)comps[comp].n_mol);
      double r = or_rand_uniform();
      or_trace tr;
      memset(&tr, 0, sizeof(tr));
      tr.cycle = cyc; tr.step = (int)s; tr.comp = comp; tr.mol = mol; tr.move = OR_NOMOVE;
      if (r < mc->p_translation[comp]) {
        tr.move = OR_TRANSLATION;
        if (comps[comp].n_mol > 0) single_body(&x, comp, mol, OR_TRANSLATION, &tr);
      } else if (r < mc->p_rotation[comp]) {
        tr.move = OR_ROTATION;
        if (comps[comp].n_mol > 0) single_body(&x, comp, mol, OR_ROTATION, &tr);
      } else if (r < mc->p_widom[comp]) {
        tr.move = OR_WIDOM;
        cbmc_move(&x, comp, mol, OR_WIDOM, &tr);
      } else if (r < mc->p_reinsertion[comp]) {
        tr.move = OR_REINSERTION;
        if (comps[comp].n_mol > 0) reinsertion_move(&x, comp, mol, &tr);
      } else if (r < mc->p_identity[comp]) {
        tr.move = OR_IDENTITY_SWAP;
        identity_swap_move(&x, &tr);
      } else if (r < mc->p_swap[comp]) {
        if (or_rand_uniform() < 0.5) {
          tr.move = OR_INSERTION;
          cbmc_move(&x, comp, mol, OR_INSERTION, &tr);
        } else {
          tr.move = OR_DELETION;
          if (comps[comp].n_mol > 0) cbmc_move(&x, comp, mol, OR_DELETION, &tr);
        }
      }
      if (nt < max_trace) trace[nt] = tr;
      nt++;
    }
    /* GatherStatisticsDuringSimulation (axpy.cu:378-400) */
    if (cyc % 500 == 0) {
      for (int c = 1; c < ncomp; c++) {
        if (mc->trans_total[c] > 0) update_max(mc->max_translation[c], &mc->trans_acc[c], &mc->trans_total[c], 5.0);
        if (mc->rot_total[c] > 0) update_max(mc->max_rotation[c], &mc->rot_acc[c], &mc->rot_total[c], 3.14);
      }
    }
  }
  free(x.sk_tmp);
  return nt;
}

/* CreateMolecule_InOneBox (axpy.cu:300-377) for one component: CBMC
 * insertions in the reference's draw order, each accepted when its
 * construction succeeds and Pacc > 1e-100 (InsertionMove::CreateMolecule,
 * move_struct.h:96-106, draws no acceptance uniform), repeated until n
 * molecules are created or max_attempts insertions were tried.  Returns the
 * number created (or -1 on error); the energy change is added to *de and to
 * the running energy. */
int or_mc_create_molecules(or_mc* mc, const or_box* b, or_atoms* comps, int ncomp, double* sk_ads, double* sk_fw,
                           int comp, int n, int64_t max_attempts, or_energy* de)
{
  if (!g_pool || comp < 1 || comp >= ncomp || mc->flexible[comp] || b->ntypes > OR_MAX_PSEUDO) return -1;
  mc_ctx x;
  memset(&x, 0, sizeof(x));
  x.b = b; x.comps = comps; x.ncomp = ncomp; x.mc = mc; x.sk_ads = sk_ads; x.sk_fw = sk_fw; x.create = 1;
  x.sk_tmp = (double*)calloc((size_t)2 * or_nvec(b), sizeof(double));
  int created = 0;
  for (int64_t a = 0; a < max_attempts && created < n; a++) {
    or_trace tr;
    memset(&tr, 0, sizeof(tr));
    const int mol = created > 0 ? created - 1 : 0;      /* TempVal.molecule (axpy.cu:349) */
    cbmc_move(&x, comp, mol, OR_INSERTION, &tr);
    if (tr.accepted) {
      created++;
      if (de) energy_add(de, &tr.de);
    }
  }
  free(x.sk_tmp);
  return created;
}

/* Batched form of the same driver: one box, the uniforms of move t from the
 * Philox stream (seed, stream, t) with the draw layout of mc_ctx (D_*), the
 * component drawn uniformly over the species with SwapProbability > 1e-10
 * (the reference's rejection loop over [0, ncomp) at axpy.cu:94-97 has the
 * same distribution).  *move_index is the box's move counter and advances by
 * one per step.  Cycle numbering, steps per cycle and the max-step update
 * follow or_mc_cycles. */
int or_batch_cycles(or_mc* mc, const or_box* b, or_atoms* comps, int ncomp,
                    double* sk_ads, double* sk_fw, uint64_t seed, uint64_t stream, int64_t* move_index,
                    int first_cycle, int n_cycles, or_trace* trace, int max_trace)
{
  mc_ctx x;
  memset(&x, 0, sizeof(x));
  x.b = b; x.comps = comps; x.ncomp = ncomp; x.mc = mc; x.sk_ads = sk_ads; x.sk_fw = sk_fw;
  x.philox = 1; x.seed = seed; x.stream = stream;
  if (b->ntypes > OR_MAX_PSEUDO) return -1;
  int valid[8], n_valid = 0;
  for (int c = 1; c < ncomp && c < 8; c++) {
    if (mc->p_swap[c] > 1e-10) valid[n_valid++] = c;
    /* flexible chains: at most 4 beads (no intramolecular pair beyond 1-4),
     * uncharged, grown by insertion / deletion / Widom only */
    if (mc->flexible[c] && (comps[c].molsize > 4 || mc->has_charge[c] || mc->p_reinsertion[c] > mc->p_widom[c] ||
                            mc->p_identity[c] > mc->p_reinsertion[c]))
      return -1;
  }
  if (n_valid == 0) return -1;
  x.sk_tmp = (double*)calloc((size_t)2 * or_nvec(b), sizeof(double));
  int nt = 0;
  for (int cyc = first_cycle; cyc < first_cycle + n_cycles; cyc++) {
    int64_t steps = 20;
    if (steps < mc->total_molecules) steps = mc->total_molecules;
    if (mc->use_max_step && steps > mc->max_step_per_cycle) steps = mc->max_step_per_cycle;
    for (int64_t s = 0; s < steps; s++) {
      x.move = (uint64_t)(*move_index);
      int vi = (int)(mc_draw(&x, D_COMP) * (double)n_valid);
      if (vi >= n_valid) vi = n_valid - 1;
      int comp = valid[vi];
      int mol = (int)(mc_draw(&x, D_MOL) * (double)comps[comp].n_mol);
      double r = mc_draw(&x, D_MOVE);
      or_trace tr;
      memset(&tr, 0, sizeof(tr));
      tr.cycle = cyc; tr.step = (int)s; tr.comp = comp; tr.mol = mol; tr.move = OR_NOMOVE;
      if (r < mc->p_translation[comp]) {
        tr.move = OR_TRANSLATION;
        if (comps[comp].n_mol > 0) single_body(&x, comp, mol, OR_TRANSLATION, &tr);
      } else if (r < mc->p_rotation[comp]) {
        tr.move = OR_ROTATION;
        if (comps[comp].n_mol > 0) single_body(&x, comp, mol, OR_ROTATION, &tr);
      } else if (r < mc->p_widom[comp]) {
        tr.move = OR_WIDOM;
        cbmc_move(&x, comp, mol, OR_WIDOM, &tr);
      } else if (r < mc->p_reinsertion[comp]) {
        tr.move = OR_REINSERTION;
        if (comps[comp].n_mol > 0) reinsertion_move(&x, comp, mol, &tr);
      } else if (r < mc->p_identity[comp]) {
        tr.move = OR_IDENTITY_SWAP;
        identity_swap_move(&x, &tr);
      } else if (r < mc->p_swap[comp]) {
        if (mc_draw(&x, D_COIN) < 0.5) {
          tr.move = OR_INSERTION;
          cbmc_move(&x, comp, mol, OR_INSERTION, &tr);
        } else {
          tr.move = OR_DELETION;
          if (comps[comp].n_mol > 0) cbmc_move(&x, comp, mol, OR_DELETION, &tr);
        }
      }
      if (nt < max_trace) trace[nt] = tr;
      nt++;
      (*move_index)++;
    }
    if (cyc % 500 == 0) {
      for (int c = 1; c < ncomp; c++) {
        if (mc->trans_total[c] > 0) update_max(mc->max_translation[c], &mc->trans_acc[c], &mc->trans_total[c], 5.0);
        if (mc->rot_total[c] > 0) update_max(mc->max_rotation[c], &mc->rot_acc[c], &mc->rot_total[c], 3.14);
      }
    }
  }
  free(x.sk_tmp);
  return nt;
}
