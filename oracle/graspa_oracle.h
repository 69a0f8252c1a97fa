/*
 * graspa_oracle.h -- FP64 CPU restatement of the gRASPA per-move GCMC energy path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine in
 * graspa_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product never links, imports or falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the gRASPA source tree, src_clean/).  Internal units are the reference's:
 * energy 10 J/mol, length Angstrom, charge e.
 *
 * Pinning (each function is checked against at least one of these in tests/):
 *   - NIST SPC/E known answers, 4 triclinic boxes (LJ, tail, real, Fourier,
 *     self, intra, k-vector count)                  tests/test_oracle_fixtures.py
 *   - gRASPA's printed energies for the TIP4P restart deck and the charged
 *     MFI-2x2x2 framework (framework S(k), self, intra; CO2 template exclusion)
 *   - closed forms: LJ(sigma)=0, LJ(2^(1/6) sigma)=-eps, NaCl Madelung constant
 *     through the Ewald sum, brute-force lattice sums vs Ewald, ideal-gas Widom
 *     weight W=1, flat-potential Rosenbluth average                tests/test_oracle_closed_forms.py
 *   - Philox4x32-10 published known-answer vectors (Salmon et al., SC'11)
 *   - glibc rand(): the oracle calls the C library's own rand() (the reference
 *     generator itself); the engine's restatement is checked against it.
 * Functions without such a check are marked "parity unpinned" at their
 * definition.
 */
#ifndef GRASPA_ORACLE_H
#define GRASPA_ORACLE_H

#include <stdint.h>

#define OR_MAX_PSEUDO 32   /* pseudo-atom types of a box */

#ifdef __cplusplus
extern "C" {
#endif

/* Simulation box + force field.  Cell layout follows read_data.cpp:1545-1547:
 * a = cell[0..2], b = cell[3..5], c = cell[6..8] (column vectors). */
typedef struct {
  double cell[9];
  double inv_cell[9];
  int    cubic;
  double volume;
  int    ntypes;
  const double* eps;      /* [ntypes*ntypes] internal energy units          */
  const double* sig;      /* [ntypes*ntypes] Angstrom                        */
  const double* shift;    /* [ntypes*ntypes] energy shift at the cutoff      */
  const double* tail;     /* [ntypes*ntypes] tail coefficient (0 if unused)  */
  const int*    use_tail; /* [ntypes*ntypes]                                 */
  double cut_vdw2;        /* squared cutoffs (read_data.cpp:522-523)         */
  double cut_coul2;
  double overlap;         /* OverlapCriteria                                 */
  int    no_charges;      /* ChargeMethod None                               */
  double alpha;
  double prefactor;       /* 138935.483496                                   */
  int    kmax[3];
  double recip_cut;
  int    lammps_ewald;
} or_box;

/* One component's atom store.  Component 0 is the framework (one molecule,
 * molsize = atom count).  Molecule m occupies atoms [m*molsize, (m+1)*molsize). */
typedef struct {
  int      molsize;
  int      n_mol;
  int      capacity;      /* molecules that fit                              */
  double*  pos;           /* [capacity*molsize*3]                            */
  double*  charge;        /* [capacity*molsize]                              */
  int*     type;          /* [capacity*molsize]                              */
  int64_t* molid;         /* [capacity*molsize]                              */
} or_atoms;

/* MoveEnergy (data_struct.h:416-442), the fields on the path. */
typedef struct {
  double hh_vdw, hg_vdw, gg_vdw;
  double hh_real, hg_real, gg_real;
  double hh_ewald, hg_ewald, gg_ewald;
  double tail;
} or_energy;

/* ---- box geometry ---- */
void   or_box_finalize(or_box* b);
void   or_pbc(double* d3, const double* cell, const double* inv, int cubic);
void   or_vdw(const or_box* b, int ta, int tb, double rr, double scaling, double out[2]);
double or_coulomb_real(const or_box* b, double qa, double qb, double r, double scaling);
int    or_nvec(const or_box* b);

/* ---- totals ---- */
void or_total_vdw_real(const or_box* b, const or_atoms* comps, int ncomp, or_energy* e);
/* Returns raw Fourier sums in e (hh/hg/gg as computed before the self and
 * exclusion terms), the stored structure factors (interleaved re,im, length
 * 2*nvec) and per-component self / intra-exclusion energies. */
void or_ewald_total(const or_box* b, const or_atoms* comps, int ncomp,
                    or_energy* fourier, double* sk_fw, double* sk_ads,
                    double* self_per_comp, double* intra_per_comp, int* n_active);
void or_exclusion_rigid(const or_box* b, const or_atoms* comp, double* intra, double* self);
double or_tail_total(const or_box* b, const int64_t* n_pseudo);
double or_tail_delta(const or_box* b, const int64_t* n_pseudo, const int64_t* d_pseudo, int sign);

/* ---- per-move primitives ---- */
/* Energies of n_sites probe sites against every atom of every component,
 * skipping atoms whose (component, molid) equals (self_comp, self_molid) or
 * (excl_comp, excl_molid).  out4[s*4+{0,1,2,3}] = hg_vdw, hg_real, gg_vdw,
 * gg_real.  overlap[s] = 1 if any pair of site s exceeded OverlapCriteria or
 * came closer than 0.1 A. */
void or_site_energies(const or_box* b, const or_atoms* comps, int ncomp,
                      int n_sites, const double* site_pos, const int* site_type,
                      const double* site_charge,
                      int self_comp, int64_t self_molid, int excl_comp, int64_t excl_molid,
                      double* out4, uint8_t* overlap);
/* Fourier-space energy change.  Returns raw SameSum (before the rigid
 * exclusion correction) and CrossSum (before the factor 2); writes the
 * updated same-type structure factor into sk_new (may be NULL). */
void or_ewald_delta(const or_box* b, const double* sk_same, const double* sk_cross,
                    int n_old, const double* old_pos, const double* old_q,
                    int n_new, const double* new_pos, const double* new_q,
                    double* same, double* cross, double* sk_new);

/* ---- proposals ---- */
/* Rotate_Quaternions (mc_utilities.h:423-457): v rotated by the random
 * rotation of the three uniforms (u, vv, w). */
void   or_rotate_quaternion(double v[3], double u, double vv, double w);

/* ---- random numbers ---- */
double or_rand_uniform(void);               /* rand()/RAND_MAX, data_struct.cpp:10-15 */
void   or_srand(unsigned seed);
void   or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double or_philox_uniform(uint64_t seed, uint64_t stream, uint64_t index, uint32_t draw);

/* ---- Monte Carlo driver in reference draw order (parity mode) ---- */
enum { OR_TRANSLATION = 0, OR_ROTATION = 1, OR_INSERTION = 2, OR_DELETION = 3, OR_WIDOM = 4,
       OR_NOMOVE = 5, OR_REINSERTION = 6, OR_IDENTITY_SWAP = 7 };

typedef struct {
  int     cycle, step, move, comp, mol, accepted, overlap, _pad;
  double  pacc;
  double  rosen;
  or_energy de;
} or_trace;

typedef struct {
  /* per adsorbate component (index c >= 1 in the component array) */
  double  p_translation[8], p_rotation[8], p_widom[8], p_swap[8];   /* cumulative, normalised */
  double  mol_fraction[8], fugacity_coeff[8], ideal_rosen[8];
  double  excl_intra[8], excl_self[8];
  double  max_translation[8][3], max_rotation[8][3];
  int64_t pseudo_per_mol[8][OR_MAX_PSEUDO];     /* pseudo-atom counts per molecule, by type */
  int     has_charge[8];
  int     n_trial_pos, n_trial_orient;
  int     vdw_real_bias;
  int     use_max_step, max_step_per_cycle;
  int     has_tail;
  double  temperature, pressure_pa, beta, pressure;
  unsigned seed;
  int     pool_size;
  int     production;                 /* 1 = record Widom statistics */
  /* running state */
  int64_t n_pseudo[OR_MAX_PSEUDO];
  int64_t total_molecules;            /* includes the framework molecule */
  int64_t trans_total[8], trans_acc[8], rot_total[8], rot_acc[8];
  int64_t ins_total[8], ins_acc[8], del_total[8], del_acc[8], widom_total[8];
  double  widom_rosen_sum[8], widom_rosen_n[8];
  or_energy running;
  /* reinsertion (ReinsertionMove, move_struct.h:170-407): cumulative
   * probability between p_widom and p_swap, and its counters */
  double  p_reinsertion[8];
  int64_t reins_total[8], reins_acc[8];
  /* identity swap (IdentitySwapMove, mc_swap_moves.h:199-426): cumulative
   * probability between p_reinsertion and p_swap; attempts and acceptances
   * from component [old] to component [new] (IdentitySwap_Total_TO / _Acc_TO) */
  double  p_identity[8];
  int64_t idsw_total[8][8], idsw_acc[8][8];
  /* flexible linear chains (no reference counterpart: the reference accepts
   * rigid molecules only, read_data.cpp:2104-2116).  Bead i >= 1 is bonded to
   * bead i - 1 at bond[c][i]; the angle (i-2, i-1, i) has the harmonic bend
   * U = bend_k/2 (theta - bend_theta0)^2 (i >= 2); the dihedral (i-3 .. i)
   * has U = t0 + t1 (1 + cos phi) + t2 (1 - cos 2 phi) + t3 (1 + cos 3 phi)
   * with phi = 180 degrees trans (i >= 3); energies in internal units.
   * Batched form only. */
  int     flexible[8];
  double  bond[8][8], bend_theta0[8][8], bend_k[8][8], torsion[8][8][4];
} or_mc;

/* Allocates the device-style random pool and seeds it (data_struct.h:1287-1346).
 * Must be called once before or_mc_cycles. */
int  or_mc_setup(or_mc* mc);
void or_mc_free(void);
/* Runs n_cycles cycles of RunMoves (axpy.cu:102-298, 580-641) and writes one
 * trace record per step into trace (capacity max_trace); returns the number of
 * records written or -1 on error.  sk_ads / sk_fw are the running structure
 * factors (2*nvec doubles each) and are updated on accepted moves. */
int  or_mc_cycles(or_mc* mc, const or_box* b, or_atoms* comps, int ncomp,
                  double* sk_ads, double* sk_fw, int first_cycle, int n_cycles,
                  or_trace* trace, int max_trace);

/* CreateMolecule_InOneBox (axpy.cu:300-377) for one component: up to n CBMC
 * insertions accepted whenever their construction succeeds (no acceptance
 * uniform drawn), at most max_attempts tries; returns the number created. */
int  or_mc_create_molecules(or_mc* mc, const or_box* b, or_atoms* comps, int ncomp, double* sk_ads, double* sk_fw,
                            int comp, int n, int64_t max_attempts, or_energy* de);

/* Batched form of the driver (the engine's device-resident GCMC): move t of
 * the box draws uniform d from or_philox_uniform(seed, stream, t, d) --
 * d0 component, d1 molecule, d2 move, d3 insertion/deletion coin, d4
 * Metropolis, d5 first-bead selection, d6 orientation selection, d8..d10
 * translation / rotation, d16+3k.. trial k, d16+3K+3o.. orientation o,
 * d16+3K+3O+3o.. retrace orientation o of a reinsertion or identity swap;
 * an identity swap takes its old species from d11 (uniform over the species
 * holding molecules), its new species from d12 (uniform over all) and the
 * old molecule from d13. */
int  or_batch_cycles(or_mc* mc, const or_box* b, or_atoms* comps, int ncomp,
                     double* sk_ads, double* sk_fw, uint64_t seed, uint64_t stream, int64_t* move_index,
                     int first_cycle, int n_cycles, or_trace* trace, int max_trace);

/* One trial position of bead i (1 <= i < molsize <= 4) of flexible component
 * comp given beads 0 .. i-1 in prev, from the batched form's draws of trial j
 * of move `move` of stream (seed, stream): bead 1 uniform on the bond
 * sphere, bead 2 with its bend angle from sin(theta) exp(-beta U_bend) and a
 * uniform azimuth, bead 3 with bend angle and dihedral from their Boltzmann
 * densities (the two are independent given beads 0..2). */
void or_flex_trial(const or_mc* mc, int comp, int i, const double* prev, uint64_t seed, uint64_t stream,
                   uint64_t move, int j, double out[3]);

/* ---- batched Widom with per-insertion counter-based streams ---- */
typedef struct {
  double rosen;          /* Rosenbluth weight W of the insertion          */
  double energy_total;   /* dE of the selected configuration (all terms)  */
  double vdw, real, ewald;
  int    first_bead, orientation, ok, _pad;
} or_widom_out;

/* n_insert Widom insertions of component `comp` (templates from its first
 * molecule slot), trial k of insertion i drawn from
 * or_philox_uniform(seed, stream, i, d).  Ewald uses sk_ads / sk_fw. */
int or_widom_batch(const or_box* b, const or_atoms* comps, int ncomp, int comp,
                   const double* sk_ads, const double* sk_fw,
                   double beta, int n_trial_pos, int n_trial_orient, int vdw_real_bias,
                   double excl_intra, double excl_self, double tail_de,
                   uint64_t seed, uint64_t stream, int64_t first_index, int n_insert,
                   or_widom_out* out);

/* Threads of the OpenMP loop over insertions in or_widom_batch; n < 1 keeps
 * the current setting.  Returns the number of threads the loop will use. */
int or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
