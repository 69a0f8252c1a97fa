/*
 * graspa_engine.h -- C-ABI of the MI355X GCMC energy engine.
 *
 * The engine owns one simulation box resident in HBM: framework atoms, guest
 * molecules, the force-field table, the Ewald tables and the running structure
 * factors.  The per-move entry points below are the calls the gRASPA Monte
 * Carlo driver makes into its energy code; each cites the reference interface
 * it replaces (gRASPA src_clean/ file:line).  The reference has no FFI for this
 * path: its boundary is the set of in-process energy calls made from one host
 * thread on the default stream, and this header is that boundary as plain C.
 *
 * Conventions
 *   - Internal units of the reference: energy 10 J/mol, length Angstrom, charge e.
 *   - All pointers are host pointers unless a name says otherwise; arrays of
 *     positions are xyz-interleaved doubles.
 *   - Every function returns GRASPA_OK (0) or a negative GRASPA_ERR_* code and
 *     never exits the process (the reference prints and calls exit(); see
 *     VDW_Coulomb.cuh:58-66).  graspa_last_error() returns the message of the
 *     last failure on the calling thread.
 *   - One host thread per box.  Calls on one box are ordered on the box's own
 *     HIP stream and are synchronous unless documented otherwise.
 *   - Component 0 is the framework; adsorbate components are numbered from 1 in
 *     the order graspa_box_add_component returns them.
 */
#ifndef GRASPA_ENGINE_H
#define GRASPA_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRASPA_OK                0
#define GRASPA_ERR_ARG          -1
#define GRASPA_ERR_HIP          -2
#define GRASPA_ERR_CAPACITY     -3
#define GRASPA_ERR_STATE        -4
#define GRASPA_ERR_UNSUPPORTED  -5

#define GRASPA_MAX_PSEUDO       32  /* pseudo-atom types a box may have */

#define GRASPA_FP64   0   /* plain FP64 form: pair terms, Ewald and sums in FP64           */
#define GRASPA_MIXED  1   /* fast form: FP64 pair arithmetic with an FP32-seeded reciprocal and a tabulated
                             erfc(ar)/r, FP32 overlap screen; FP64 accumulation, Ewald and S(k) */

typedef struct graspa_box graspa_box;
typedef struct graspa_mc graspa_mc;

/* Box, force field and Ewald set-up.  Replaces the state the reference keeps in
 * Boxsize / ForceField (data_struct.h:838-886) after read_data.cpp's parsers. */
typedef struct {
  double cell[9];            /* a = cell[0..2], b = cell[3..5], c = cell[6..8] (read_data.cpp:1545-1547) */
  int    ntypes;             /* pseudo-atom types                                          */
  const double* epsilon;     /* [ntypes*ntypes] mixed epsilon, internal units               */
  const double* sigma;       /* [ntypes*ntypes] mixed sigma, Angstrom                       */
  const double* shift;       /* [ntypes*ntypes] energy shift at the cutoff (0 = truncated)  */
  double cutoff_vdw;         /* Angstrom                                                    */
  double cutoff_coulomb;     /* Angstrom                                                    */
  double overlap_criteria;   /* OverlapCriteria (read_data.cpp:455)                        */
  int    use_charges;        /* ChargeMethod Ewald = 1, None = 0                            */
  double alpha;              /* Ewald alpha, 1/Angstrom                                     */
  int    kmax[3];
  double reciprocal_cutoff;  /* integer k^2 cut, or |k|^2 cut when lammps_ewald            */
  int    lammps_ewald;       /* Ewald_UseLAMMPS_Setup                                       */
  double prefactor;          /* Coulomb prefactor 138935.483496 (read_data.cpp:611)        */
  int    precision;          /* GRASPA_FP64 or GRASPA_MIXED                                 */
  int    device;             /* HIP device ordinal                                          */
} graspa_box_params;

/* MoveEnergy (data_struct.h:416-442): the fields on the GCMC energy path. */
typedef struct {
  double hh_vdw, hg_vdw, gg_vdw;
  double hh_real, hg_real, gg_real;
  double hh_ewald, hg_ewald, gg_ewald;
  double tail;
} graspa_energy;

/* Per-insertion result of a batched Widom call. */
typedef struct {
  double rosen;              /* Rosenbluth weight W (first bead x chain x Ewald x tail)      */
  double energy;             /* dE of the selected configuration, all terms                 */
  double vdw, real, ewald;   /* its parts                                                   */
  int    first_bead;         /* selected first-bead trial                                   */
  int    orientation;        /* selected orientation                                        */
  int    ok;                 /* 0 if the CBMC construction failed (W = 0)                   */
  int    _pad;
} graspa_widom_result;

/* Sums over a batched Widom call (device-side reduction). */
typedef struct {
  double sum_w;              /* sum of W                                                    */
  double sum_w2;             /* sum of W^2                                                  */
  double sum_we;             /* sum of W * dE                                               */
  double n;                  /* insertions                                                  */
  double n_first_bead;       /* insertions whose first bead was placed (orientations grown) */
  double n_ok;               /* insertions with a complete construction (W > 0)             */
  double screened_pairs;     /* work: (trial site, atom) pairs tested by the overlap screen */
  double evaluated_pairs;    /* work: (trial site, atom) pairs with full energies           */
} graspa_widom_summary;

/* ------------------------------------------------------------------------ */
/* Box life cycle and state                                                  */
/* ------------------------------------------------------------------------ */
const char* graspa_last_error(void);
const char* graspa_version(void);
int  graspa_device_count(int* n);

/* Creates a box on params->device.  Replaces the allocation part of
 * Initialize() (main.cpp:58-343) and Copy_ForceField_to_GPU (fxn_main.h:129-150). */
int  graspa_box_create(const graspa_box_params* params, graspa_box** out);
void graspa_box_destroy(graspa_box* box);

/* Framework atoms (component 0, one rigid molecule): CheckFrameworkCIF output
 * (read_data.cpp:1474-1849). */
int  graspa_box_set_framework(graspa_box* box, int n_atoms, const double* xyz, const double* charge,
                              const int* type);

/* Adds an adsorbate species with room for capacity_molecules molecules.  The
 * template (MoleculeDefinitionParser, read_data.cpp:2044-2195) fills slot 0
 * until a molecule is stored there, as in the reference.  *comp_out is >= 1. */
int  graspa_box_add_component(graspa_box* box, int molsize, int capacity_molecules, const double* template_xyz,
                              const int* template_type, const double* template_charge, int* comp_out);

/* Device memory the box holds now (framework copies and screen cells, guest
 * store, tables, structure factors, voxel maps, Widom buffers grown so far),
 * in bytes.  No counterpart in the reference; a job server admits boxes by
 * it.  A batch holding the box adds its own per-slot arrays. */
int  graspa_box_footprint(graspa_box* box, uint64_t* bytes);

/* Bonded model of a flexible linear chain of at most 4 beads, grown bead by
 * bead in CBMC moves.  No counterpart in the reference, which grows rigid
 * molecules only (read_data.cpp:2104-2116).  Entry i belongs to bead i. */
typedef struct {
  double bond[4];          /* length of the bond to bead i - 1 (i >= 1), Angstrom                    */
  double theta0[4];        /* equilibrium angle (i-2, i-1, i) (i >= 2), rad                           */
  double k_bend[4];        /* U = k_bend/2 (theta - theta0)^2, internal energy units per rad^2        */
  double torsion[4][4];    /* dihedral (i-3 .. i) (i >= 3): t0 + t1 (1 + cos phi) + t2 (1 - cos 2 phi)
                              + t3 (1 + cos 3 phi), phi = pi trans, internal energy units            */
} graspa_chain;
/* Makes component comp a flexible chain (chain != NULL) or rigid again
 * (NULL).  A flexible component has at most 4 beads (no intramolecular pair
 * beyond 1-4) and no charges.  Its insertions, deletions and Widom moves in
 * the batched driver (graspa_batch_*) grow bead i from n_trial_orient
 * positions drawn from the bend and dihedral Boltzmann densities given the
 * beads placed (deletion: trial 0 is the molecule's own bead) and select one
 * on the Boltzmann factor of its external energy; W is the product of the
 * first bead's and each step's trial averages.  Translation and rotation move
 * the chain rigidly.  The parity driver (graspa_mc_*), the Widom pipeline and
 * reinsertion / identity-swap moves refuse flexible components. */
int  graspa_box_set_chain(graspa_box* box, int comp, const graspa_chain* chain);

/* Replaces all molecules of a component (RestartFileParser, read_data.cpp:3000-3221). */
int  graspa_box_set_molecules(graspa_box* box, int comp, int n_mol, const double* xyz, const double* charge,
                              const int* type, const int64_t* molid);
/* Copies a component's atoms back to the host; arrays sized for n_mol*molsize
 * atoms (any pointer may be NULL). */
int  graspa_box_get_molecules(graspa_box* box, int comp, int* n_mol, double* xyz, double* charge, int* type,
                              int64_t* molid);

/* Builds the framework and adsorbate structure factors on the device from the
 * current atoms and returns the raw Fourier sums (Ewald_Total,
 * ewald_preparation.h:5-259; the stored guest S(k) is total minus framework). */
int  graspa_box_init_ewald(graspa_box* box, graspa_energy* fourier);
/* Structure factors as 2*nvec doubles (re,im) in the reference's kxyz order
 * (Ewald_Energy_Functions.h:296-300).  Either pointer may be NULL. */
int  graspa_box_get_structure_factors(graspa_box* box, double* sk_framework, double* sk_adsorbate);
int  graspa_box_set_structure_factors(graspa_box* box, const double* sk_framework, const double* sk_adsorbate);
int  graspa_box_nvec(graspa_box* box, int* nvec, int* n_active);

/* ------------------------------------------------------------------------ */
/* Per-move energies (the hot path)                                          */
/* ------------------------------------------------------------------------ */

/* Pair energies of n_sites probe sites against every framework and guest atom,
 * skipping atoms of molecule (self_comp, self_molid) and of the ExcludeList
 * entry (excl_comp, excl_molid); pass -1 to disable either.
 * out4[4*s+{0,1,2,3}] = host-guest VDW, host-guest real, guest-guest VDW,
 * guest-guest real.  overlap[s] = 1 if a pair of site s exceeded
 * OverlapCriteria or came within 0.1 A.
 * Replaces Calculate_Multiple_Trial_Energy_VDWReal (VDW_Coulomb.cuh:93,
 * VDW_Coulomb.cu:1183-1352) with its per-block partials summed on the device
 * instead of in Host_sum_Widom_HGGG_SEPARATE (mc_widom.h:41-87). */
int  graspa_site_energies(graspa_box* box, int n_sites, const double* xyz, const int* type, const double* charge,
                          int self_comp, int64_t self_molid, int excl_comp, int64_t excl_molid,
                          double* out4, uint8_t* overlap);

/* CBMC trial batch: n_trials trials of `chain` sites each (site s of trial t at
 * xyz[3*(t*chain+s)]; types and charges are per chain position).  Per-trial
 * sums out4[4*t+...] and per-trial overlap flags, as CBMC_PairwiseInteractions
 * (mc_widom.h:89-119) returns them. */
int  graspa_trial_energies(graspa_box* box, int n_trials, int chain, const double* xyz, const int* type,
                           const double* charge, int self_comp, int64_t self_molid, int excl_comp,
                           int64_t excl_molid, double* out4, uint8_t* overlap);

/* Energy change of moving molecule (comp, molid) of `molsize` atoms from old_xyz
 * to new_xyz (HG/GG x VDW/real; Ewald fields left 0); *overlap from the new
 * position only.  Replaces Calculate_Single_Body_Energy_VDWReal
 * (VDW_Coulomb.cuh:91, VDW_Coulomb.cu:626-841) with the host block sum of
 * SingleBody_Calculation (mc_single_particle.h:183-200). */
int  graspa_single_body_delta(graspa_box* box, int comp, int64_t molid, int molsize, const double* old_xyz,
                              const double* new_xyz, const int* type, const double* charge, graspa_energy* de,
                              int* overlap);

/* Fourier-space energy change of removing n_old charges and adding n_new
 * (same-type = adsorbate S(k), cross = framework S(k)).  *same is the raw
 * same-type sum and *cross the raw cross sum: the reference's
 * GPU_EwaldDifference_General (VDW_Coulomb.cuh:112, Ewald_Energy_Functions.h:438-580)
 * returns {same - exclusion, 2*cross}.  The updated adsorbate S(k) is staged on
 * the device until graspa_ewald_accept. */
int  graspa_ewald_delta(graspa_box* box, int n_old, const double* old_xyz, const double* old_charge, int n_new,
                        const double* new_xyz, const double* new_charge, double* same, double* cross);
/* Makes the staged S(k) current: Update_Vector_Ewald (Ewald_Energy_Functions.h:423-433). */
int  graspa_ewald_accept(graspa_box* box);

/* Commits.  update_translation_position (mc_utilities.h:285-292),
 * Update_insertion_data_Parallel (mc_utilities.h:191-233) and
 * Update_deletion_data_Parallel (mc_utilities.h:163-189): a deletion copies the
 * last molecule's positions, charges and types into the hole, whose MolIDs stay. */
int  graspa_commit_move(graspa_box* box, int comp, int mol_slot, const double* new_xyz);
int  graspa_commit_insert(graspa_box* box, int comp, const double* xyz, const double* charge, const int* type,
                          int64_t molid);
int  graspa_commit_delete(graspa_box* box, int comp, int mol_slot);

/* Total VDW + real-space energies (HG, GG) of the current configuration and
 * raw Fourier sums from the running S(k): Total_VDW_Coulomb_Energy
 * (VDW_Coulomb.cu:1580-1665) and Ewald_TotalEnergy (Ewald_Energy_Functions.h:1272-1430). */
int  graspa_total_energy(graspa_box* box, graspa_energy* e);

/* ------------------------------------------------------------------------ */
/* Batched Widom / CBMC insertions                                           */
/* ------------------------------------------------------------------------ */

/* n_insert independent Widom test insertions of component `comp`, each a full
 * CBMC construction (n_trial_pos first-bead trials, n_trial_orient rigid-body
 * orientations, Ewald and tail of the selected configuration), in one
 * persistent launch.  Insertion i draws its random numbers from the
 * counter-based Philox4x32-10 stream (seed, stream, first_index + i), so a
 * result does not depend on batch size or scheduling.  The semantics are
 * Insertion_Body (mc_swap_utilities.h:3-133) on the WIDOM branch of RunMoves
 * (axpy.cu:163-186).  excl = ExclusionIntra + ExclusionAtom of the species
 * (ewald_preparation.h:351-366).  results may be NULL (summary only). */
int  graspa_widom_batch(graspa_box* box, int comp, int n_trial_pos, int n_trial_orient, double beta,
                        int vdw_real_bias, double excl, double tail_de, uint64_t seed, uint32_t stream,
                        int64_t first_index, int n_insert, graspa_widom_result* results,
                        graspa_widom_summary* summary);

/* ------------------------------------------------------------------------ */
/* GCMC driver in the reference's random-number order (parity mode)          */
/* ------------------------------------------------------------------------ */

/* Per-species move set-up, the Components fields RunMoves reads. */
typedef struct {
  double p_translation, p_rotation, p_widom, p_swap;   /* cumulative, normalised (data_struct.h:569-608) */
  double mol_fraction, fugacity_coeff, ideal_rosen;
  double excl_intra, excl_self;                        /* rigid exclusion (ewald_preparation.h:351-366) */
  double max_translation[3], max_rotation[3];
  int64_t pseudo_per_mol[GRASPA_MAX_PSEUDO];                          /* pseudo-atom counts of one molecule            */
  double p_reinsertion;            /* cumulative, between p_widom and p_swap (a value <= p_widom: none)  */
  double p_identity;               /* cumulative, between p_reinsertion and p_swap (<= p_reinsertion: none) */
} graspa_species;

typedef struct {
  int     n_species;               /* adsorbate components 1..n_species                      */
  graspa_species species[8];
  int     n_trial_pos, n_trial_orient;
  int     vdw_real_bias;
  int     use_max_step, max_step_per_cycle;
  int     has_tail;
  const double* tail;              /* [ntypes*ntypes] tail coefficients (may be NULL)         */
  const int*    use_tail;          /* [ntypes*ntypes]                                         */
  double  beta, pressure;          /* internal units                                          */
  unsigned seed;                   /* srand seed (RandomSeed)                                  */
  int     pool_size;               /* device random pool entries (main.cpp:145: 333334)       */
  int64_t n_pseudo[GRASPA_MAX_PSEUDO];            /* pseudo-atom counts of the whole box                      */
} graspa_mc_params;

/* One record per MC step. */
typedef struct {
  int     cycle, step, move, comp, mol, accepted, overlap, _pad;
  double  pacc;
  double  rosen;
  graspa_energy de;
} graspa_trace;

#define GRASPA_MOVE_TRANSLATION 0
#define GRASPA_MOVE_ROTATION    1
#define GRASPA_MOVE_INSERTION   2
#define GRASPA_MOVE_DELETION    3
#define GRASPA_MOVE_WIDOM       4
#define GRASPA_MOVE_NONE        5
#define GRASPA_MOVE_REINSERTION 6   /* ReinsertionMove (move_struct.h:170-407) */
#define GRASPA_MOVE_IDENTITY_SWAP 7 /* IdentitySwapMove (mc_swap_moves.h:199-426); the trace's comp and
                                       mol name the old molecule                                      */
/* Move counters per species: trans_tot, trans_acc, rot_tot, rot_acc,
 * ins_tot, ins_acc, del_tot, del_acc, widom_tot, reins_tot, reins_acc, then
 * identity swaps from this species to species s (s = 0 .. 6, adsorbate
 * component s + 1) attempted at [GRASPA_CNT_IDSW_TOT + s] and accepted at
 * [GRASPA_CNT_IDSW_ACC + s] (IdentitySwap_Total_TO / _Acc_TO). */
#define GRASPA_CNT_IDSW_TOT    11
#define GRASPA_CNT_IDSW_ACC    18
#define GRASPA_N_COUNTERS      25

/* RandomNumber::Setup (data_struct.h:1287-1346) plus the move bookkeeping of
 * Initialize (main.cpp:58-343).  Uses the engine's restatement of glibc rand(). */
int  graspa_mc_create(graspa_box* box, const graspa_mc_params* params, graspa_mc** out);
void graspa_mc_destroy(graspa_mc* mc);
/* Run_Simulation_ForOneBox (axpy.cu:627-641) for n_cycles cycles starting at
 * cycle index first_cycle; production != 0 records Widom statistics.  Returns the
 * number of steps in *n_steps; up to max_trace records go to trace (may be NULL). */
int  graspa_mc_cycles(graspa_mc* mc, int first_cycle, int n_cycles, int production, graspa_trace* trace,
                      int max_trace, int64_t* n_steps);
/* CreateMolecule_InOneBox (axpy.cu:300-377) for one component: CBMC
 * insertions in the reference's draw order, each accepted when its
 * construction succeeds and Pacc > 1e-100 (InsertionMove::CreateMolecule,
 * move_struct.h:96-106, no acceptance uniform, no move counters), until n
 * molecules exist more or max_attempts insertions were tried
 * (GRASPA_ERR_STATE then, as the reference throws).  The energy change goes to
 * *de and into the running energy. */
int  graspa_mc_create_molecules(graspa_mc* mc, int comp, int n, int64_t max_attempts, int* n_created,
                                graspa_energy* de);
/* Running energy change (SystemComponents.deltaE), molecule counts, move
 * counters [species][GRASPA_N_COUNTERS], Widom sums. */
int  graspa_mc_stats(graspa_mc* mc, graspa_energy* running, int* n_mol, int64_t* counters, double* widom_sum_w,
                     double* widom_n);
/* Per-step wall-clock latency (seconds) of the last graspa_mc_cycles call,
 * one entry per cycle; returns the number written. */
int  graspa_mc_cycle_times(graspa_mc* mc, double* seconds, int max_n);

/* ------------------------------------------------------------------------ */
/* Device-resident GCMC of many boxes (batched form)                         */
/* ------------------------------------------------------------------------ */

/* A batch runs the Monte Carlo steps of many independent boxes in one
 * persistent launch, one workgroup per box: component / molecule / move
 * selection, translation and rotation proposals, CBMC trial batches with their
 * selection, the Ewald energy change, the Metropolis test, the commit and the
 * structure-factor flip all run on the device.  It replaces RunMoves
 * (axpy.cu:102-298) and Run_Simulation_ForOneBox (axpy.cu:378-400, 580-641)
 * for jobs that do not need the reference's random-number sequence: move t of
 * a box draws from the Philox4x32-10 stream (seed, stream, t), so a box's
 * trajectory is the same alone and in any batch.  Draw layout of a move:
 * d0 component (uniform over the species with SwapProbability > 0), d1
 * molecule, d2 move type, d3 insertion / deletion coin, d4 Metropolis, d5
 * first-bead selection, d6 orientation selection, d8..d10 translation or
 * rotation, d11 / d12 / d13 the old species (uniform over the species
 * holding molecules), new species (uniform over all) and old molecule of an
 * identity swap, d16+3k.. first-bead trial k, d16+3K+3o.. orientation o,
 * d16+3K+3O+3o.. retrace orientation o (reinsertion, identity swap).
 *
 * Every box of a batch is on one device and has one precision form.  Between
 * graspa_batch_create and graspa_batch_destroy the boxes belong to the batch:
 * after graspa_batch_sync their host-side state (molecules, counts, structure
 * factors) is current and the read-only box calls may be used. */
typedef struct graspa_batch graspa_batch;

typedef struct {
  graspa_box* box;
  const graspa_mc_params* params;  /* species, trials, bias, beta, pressure, tail, max step
                                      (seed and pool_size are not used)                          */
  uint64_t seed;                   /* Philox key of this box's moves                            */
  uint32_t stream;                 /* Philox stream of this box                                 */
  int      n_block;                /* blocks of a production run, 1..GRASPA_NBLOCK (0: 5)        */
} graspa_batch_box;

/* n_boxes slots; boxes[i].box == NULL (or boxes == NULL) leaves slot i empty.
 * trace_cap: move records kept per box and run (0 = none); cycle_cap:
 * per-cycle latencies kept per box (ring, 0 = none). */
int  graspa_batch_create(const graspa_batch_box* boxes, int n_boxes, int trace_cap, int cycle_cap,
                         graspa_batch** out);
void graspa_batch_destroy(graspa_batch* batch);
/* Continuous batching: puts a box into a slot (NULL empties it; the slot's
 * statistics start from zero), gives a slot n_cycles more cycles, and runs
 * one launch of at most moves_per_launch moves per box (*n_running = boxes
 * still short of their cycles).  A slot can be refilled while the others
 * keep their state. */
int  graspa_batch_set_box(graspa_batch* batch, int slot, const graspa_batch_box* box);
int  graspa_batch_set_cycles(graspa_batch* batch, int slot, int n_cycles, int production);
int  graspa_batch_launch(graspa_batch* batch, int moves_per_launch, int* n_running);
/* graspa_batch_launch in two halves, so the host can prepare boxes while the
 * device runs: launch_async queues the launch and returns; wait finishes it
 * (*n_running as above).  Other calls on the batch wait for a launch in
 * flight first. */
int  graspa_batch_launch_async(graspa_batch* batch, int moves_per_launch);
int  graspa_batch_wait(graspa_batch* batch, int* n_running);
/* Time slices for continuous batching (no counterpart in the reference):
 * with us > 0 each launch gives every box a move budget of about us
 * microseconds at its own measured time per move (device clock, updated after
 * every launch), at most moves_per_launch, so fast boxes are not held idle
 * until the slowest box of the launch ends; us = 0 (the default) gives every
 * box moves_per_launch moves.  A box's trajectory does not depend on it. */
int  graspa_batch_set_time_slice(graspa_batch* batch, double us);
/* Advances every box by n_cycles cycles (production != 0 records Widom and
 * loading statistics); launches of at most moves_per_launch moves per box
 * (<= 0: a default) until all boxes are done.  *moves = moves made. */
int  graspa_batch_run(graspa_batch* batch, int n_cycles, int production, int moves_per_launch, int64_t* moves);
/* State of box i: running energy change, molecule counts [n_species], move
 * counters [n_species][GRASPA_N_COUNTERS] (graspa_mc_stats layout), Widom sums [n_species],
 * per-cycle loading sums sum N and sum N^2 [n_species] over the production
 * cycles (*n_load_cycles), moves and cycles made since creation.  Any output
 * pointer may be NULL. */
int  graspa_batch_box_stats(graspa_batch* batch, int i, graspa_energy* running, int* n_mol, int64_t* counters,
                            double* widom_sum_w, double* widom_n, double* load_sum, double* load_sum2,
                            int64_t* n_load_cycles, int64_t* moves, int64_t* cycles);
/* Per-cycle wall-clock latency (seconds, device clock) of box i's cycles in
 * the last graspa_batch_run, oldest first; returns the number written. */
int  graspa_batch_cycle_times(graspa_batch* batch, int i, double* seconds, int max_n);
/* Move records of box i from the last graspa_batch_run (graspa_trace layout,
 * cycle and step counted from the batch's creation); returns the number. */
int  graspa_batch_trace(graspa_batch* batch, int i, graspa_trace* out, int max_n);
/* Moves and cycles made and cycle targets of every slot ([n_boxes] each, 0 for
 * empty slots); any pointer may be NULL. */
int  graspa_batch_progress(graspa_batch* batch, int64_t* moves, int64_t* cycles, int64_t* targets);
/* Device-clock seconds each slot's box has spent in launches since it was
 * put into the slot ([n_boxes], 0 for empty slots): a launch lasts as long as
 * its busiest box, so these against the kernel time show the load balance. */
int  graspa_batch_busy(graspa_batch* batch, double* seconds);
/* Summed kernel time (HIP events) and launches of the batch since creation. */
int  graspa_batch_kernel_time(graspa_batch* batch, double* total_ms, int64_t* launches);
/* Copies the device state of every box back to its graspa_box. */
int  graspa_batch_sync(graspa_batch* batch);
/* Current maximum translation and rotation steps of box i, [n_species][3]
 * each (Components::MaxTranslation / MaxRotation as adapted by
 * Update_Max_Translation / Update_Max_Rotation, mc_utilities.h:612-665, every
 * 500 cycles); either pointer may be NULL. */
int  graspa_batch_max_steps(graspa_batch* batch, int i, double* max_translation, double* max_rotation);

/* Block statistics of the current production run of a box, the data behind
 * the reference's block averages (axpy.cu:420-453, 176-185; Gather_Averages_*
 * in print_statistics.cuh:412-456, 646-665): the run's cycles are split into
 * n_block blocks of block_size cycles (n_block from graspa_batch_box, 5 by
 * default as the reference's NumberOfBlocks; the last block also takes the
 * remainder; block = min(cycle / block_size, n_block - 1)).  Per block
 * and component c (1..n_comp-1; 0 is the framework): sums over the block's
 * cycles of N_c, N_c^2, N_c N_d, dE N_c (dE = running energy change without
 * the host-host terms) and of the running energy change per term; Widom
 * sums of W and counts.  A production set_cycles / run starts a new run and
 * clears them. */
#define GRASPA_NBLOCK    16  /* most blocks per run */
#define GRASPA_MAX_COMPS 8
typedef struct {
  int64_t n_block;                                       /* blocks of this run                 */
  int64_t block_size;                                    /* cycles per block                   */
  int64_t cycles[GRASPA_NBLOCK];                         /* cycles recorded per block          */
  double  n[GRASPA_NBLOCK][GRASPA_MAX_COMPS];            /* sum N_c                            */
  double  n2[GRASPA_NBLOCK][GRASPA_MAX_COMPS];           /* sum N_c^2                          */
  double  nn[GRASPA_NBLOCK][GRASPA_MAX_COMPS][GRASPA_MAX_COMPS]; /* sum N_c N_d              */
  double  en[GRASPA_NBLOCK][GRASPA_MAX_COMPS];           /* sum dE N_c                         */
  double  e[GRASPA_NBLOCK][10];                          /* sum of the running graspa_energy   */
  double  widom_w[GRASPA_NBLOCK][GRASPA_MAX_COMPS];      /* sum W of Widom moves               */
  double  widom_n[GRASPA_NBLOCK][GRASPA_MAX_COMPS];      /* Widom moves                        */
} graspa_block_stats;
/* Asynchronous form of graspa_widom_batch (no counterpart in the reference,
 * whose Widom moves each end in a host round trip, mc_widom.h:67-119): the
 * batch is queued on the box's stream and the call returns without waiting.
 * Every batch of a box, queued or not, adds its summary to running totals
 * kept on the device; graspa_widom_totals waits for the queued batches,
 * returns the totals since the last reset and resets them when reset != 0.
 * The arguments are those of graspa_widom_batch. */
int  graspa_widom_enqueue(graspa_box* box, int comp, int n_trial_pos, int n_trial_orient, double beta,
                          int vdw_real_bias, double excl, double tail_de, uint64_t seed, uint32_t stream,
                          int64_t first_index, int n_insert);
int  graspa_widom_totals(graspa_box* box, graspa_widom_summary* totals, int reset);
/* graspa_widom_enqueue without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops that forbid hipMalloc).  The
 * pipeline's scratch (trial masks, per-insertion state, work lists, energies,
 * records, partial sums) comes from one fixed 32 MiB block per device in the
 * library's static device memory instead of the box's buffers, and no
 * certain-overlap voxel map is built (maps that earlier calls built are
 * used; without one the cell screen decides the same trials).  A batch that
 * needs more scratch than the block (about 18,000 CO2 insertions at 10 x 10
 * trials on 256 CUs) is refused with GRASPA_ERR_ARG before any GPU work.
 * Results are those of graspa_widom_enqueue: the summary goes into the box's
 * running totals (graspa_widom_totals).  Every call on a device shares the
 * block: calls on one device that overlap in time (other streams, other
 * host threads) race and must be serialised by the caller. */
int  graspa_widom_enqueue_noalloc(graspa_box* box, int comp, int n_trial_pos, int n_trial_orient, double beta,
                                  int vdw_real_bias, double excl, double tail_de, uint64_t seed, uint32_t stream,
                                  int64_t first_index, int n_insert);

/* A queued Widom batch captured as a HIP graph (no counterpart in the
 * reference, which launches each trial kernel from the host,
 * mc_widom.h:440-566).  graspa_widom_graph_create captures one
 * graspa_widom_enqueue batch with the given arguments (the box's current
 * launch shape, buffers and voxel maps, grown and built before the capture)
 * and keeps it in the box; its first insertion index lives in device memory
 * and the graph's last kernel advances it by n_insert, so replay r runs the
 * batch at first_index + r * n_insert, the same batches as enqueueing with
 * those indices.  graspa_widom_graph_launch queues n_replays replays on the
 * box's stream and returns at once; summaries go to the running totals
 * (graspa_widom_totals).  A replay is refused with GRASPA_ERR_STATE when the
 * box's arguments for the batch differ from the captured ones (molecules
 * committed, structure factors swapped, framework, buffers or launch shape
 * changed); capturing again fixes that.  A new capture replaces the old. */
int  graspa_widom_graph_create(graspa_box* box, int comp, int n_trial_pos, int n_trial_orient, double beta,
                               int vdw_real_bias, double excl, double tail_de, uint64_t seed, uint32_t stream,
                               int64_t first_index, int n_insert);
int  graspa_widom_graph_launch(graspa_box* box, int n_replays);

int  graspa_batch_block_stats(graspa_batch* batch, int i, graspa_block_stats* out);

/* Launch shape of the Widom pipeline's pair-energy kernel.  No counterpart in
 * the reference, whose trial-energy kernel has one fixed shape (128 threads
 * per block, one thread per (atom, site) pair, VDW_Coulomb.cu:1183-1352 as
 * launched at mc_widom.h:109). */
typedef struct {
  int waves;            /* waves per workgroup: 4, 8, 12 (default) or 16          */
  int per_cu;           /* workgroups per CU; 0 (default): as many as are resident */
  int framework_lds;    /* 1 (default): framework staged in LDS when it fits; 0: read from L2 */
} graspa_eval_shape;

/* Sets the box's launch shape (NULL: the default); GRASPA_ERR_ARG for a
 * shape outside the list above.  Results do not depend on the shape. */
int  graspa_widom_set_shape(graspa_box* box, const graspa_eval_shape* shape);
int  graspa_widom_get_shape(graspa_box* box, graspa_eval_shape* shape);
/* Autotuner: runs the given Widom batch (arguments as graspa_widom_batch,
 * stream 0, first index 0) in every launch shape that fits the device, reps
 * timed runs each after one untimed run, and keeps the fastest shape in the
 * box.  The first min(cap, *n_tried) shapes tried and their milliseconds per
 * batch are written to tried/ms (either may be NULL).  The running totals of
 * queued batches (graspa_widom_totals) and the last batch's summary are
 * saved before the timed runs and restored after them; the per-insertion
 * records of the last batch are overwritten. */
int  graspa_widom_autotune(graspa_box* box, int comp, int n_trial_pos, int n_trial_orient, double beta,
                           int vdw_real_bias, double excl, double tail_de, uint64_t seed, int n_insert, int reps,
                           graspa_eval_shape* tried, double* ms, int cap, int* n_tried);

/* ------------------------------------------------------------------------ */
/* Measurement                                                               */
/* ------------------------------------------------------------------------ */

/* Kernel timing with HIP events on the box's stream.  While enabled, every
 * launch of the named kernel family ("widom", "site", "ewald") is bracketed by
 * events; graspa_timing_read returns the summed elapsed milliseconds and the
 * number of launches since graspa_timing_enable. */
int  graspa_timing_enable(graspa_box* box, const char* family, int on);
int  graspa_timing_read(graspa_box* box, double* total_ms, int64_t* launches);
/* Blocks until all work queued on the box's stream is done. */
int  graspa_box_sync(graspa_box* box);

/* glibc-compatible rand() restatement used by the parity driver (exposed for
 * its test against the C library): srand(seed) then n draws of rand(). */
int  graspa_glibc_rand_sequence(unsigned seed, int n, int* out);

/* Host helper with no counterpart in the reference: the table of
 * erfc(alpha r)/r over r^2 that the fast (GRASPA_MIXED) Widom kernel uses for
 * the real-space Coulomb term (16 intervals per octave of r^2 from 2^-7, a
 * degree-7 polynomial in r^2 - r_i^2 each, 8 doubles per interval), exposed
 * so that it can be checked without a GPU.  Returns the number of doubles
 * written, the number needed when cap is too small, or -1 for bad arguments. */
int  graspa_coulomb_table(double alpha, double cut2, double* out, int cap);

/* Host helper with no counterpart in the reference: the map of the
 * framework's certain-overlap volume for probe sites of type site_type that
 * the overlap screens use (voxels of about 0.15 A, nv[3] per cell axis; bit
 * x + nv0 (y + nv1 z) set when every point of the voxel has a framework pair
 * above OverlapCriteria or closer than 0.1 A).  Uses params' cell, types,
 * force field, VDW cutoff and OverlapCriteria.  Returns the number of 32-bit
 * words written, the number needed when cap is too small, or -1. */
int  graspa_dead_map(const graspa_box_params* params, int n_fw, const double* xyz, const int* type, int site_type,
                     uint32_t* bits, int cap, int* nv);

/* Host helper with no counterpart in the reference: the companion map of
 * certainly free voxels (same voxels and bit layout as graspa_dead_map; bit
 * set when every point of the voxel is farther than the overlap distance from
 * every framework atom, with the squared distance raised by 1e-4 relative).
 * The first-bead screen skips trials there: the cell screen could not flag
 * them.  Same arguments and return values as graspa_dead_map. */
int  graspa_free_map(const graspa_box_params* params, int n_fw, const double* xyz, const int* type, int site_type,
                     uint32_t* bits, int cap, int* nv);

#ifdef __cplusplus
}
#endif
#endif
