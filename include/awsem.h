/*
 * awsem.h -- C ABI of the MI355X (gfx950) FixBackbone force path.
 *
 * Plain C99: pointers, sizes and status codes only.  Every call returns 0
 * on success or a negative AWSEM_E_* code; awsem_last_error() gives the
 * message (the reference's error->all text where one exists).  No call
 * aborts the process.
 *
 * Each entry point cites the reference interface it replaces (paths
 * relative to the reference repository root; fb.cpp = src/fix_backbone.cpp,
 * fb.h = src/fix_backbone.h, pev.cpp = src/pair_excluded_volume.cpp).
 *
 * Coordinate layouts
 *   atom layout    : x[n_atoms][3] FP64 in the caller's (LAMMPS) atom order,
 *                    mapped to residues by awsem_set_topology().
 *   residue layout : x[3][n][3] FP64, kind 0 = CA, 1 = CB (H-beta for Gly),
 *                    2 = O, then residue, then xyz.
 * Forces are ADDED into f (pre_force semantics: LAMMPS clears f before
 * modify->pre_force and the fix accumulates with +=, fb.cpp:8181-8189).
 */
#ifndef AWSEM_H
#define AWSEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define AWSEM_OK 0
#define AWSEM_E_ARG (-1)          /* invalid argument / shape                       */
#define AWSEM_E_FILE (-2)         /* missing or unreadable input file               */
#define AWSEM_E_FORMAT (-3)       /* malformed coeff / table / memory file          */
#define AWSEM_E_UNSUPPORTED (-4)  /* coeff enables a term outside this force path   */
#define AWSEM_E_TOPOLOGY (-5)     /* bad residue / chain tags (fb.cpp:1651-1696)    */
#define AWSEM_E_MISSING_ATOM (-6) /* DSSP/P_AP/helix missing atom (fb.cpp:2937 ...) */
#define AWSEM_E_DEVICE (-7)       /* HIP runtime failure                            */
#define AWSEM_E_NOMEM (-8)        /* allocation / HBM budget exceeded               */
#define AWSEM_E_NONFINITE (-9)    /* NaN/Inf coordinates or forces                  */
#define AWSEM_E_STATE (-10)       /* call out of order                              */
#define AWSEM_E_RANGE (-11)       /* FM table: a distance outside [rmin, rmax]      */
                                  /* (fb.cpp:4419-4427; the reference aborts)       */

/* energy slots: FixBackbone::EnergyTerms order (fb.h:331-332) + bond style */
enum awsem_energy_slot {
  AWSEM_ET_TOTAL = 0, AWSEM_ET_CHAIN, AWSEM_ET_SHAKE, AWSEM_ET_CHI, AWSEM_ET_RAMA,
  AWSEM_ET_VEXCLUDED, AWSEM_ET_DSSP, AWSEM_ET_PAP, AWSEM_ET_WATER, AWSEM_ET_BURIAL,
  AWSEM_ET_HELIX, AWSEM_ET_AMHGO, AWSEM_ET_FRAGMEM, AWSEM_ET_VFRAGMEM,
  AWSEM_ET_CONT_REST, AWSEM_ET_MEMB, AWSEM_ET_SSB, AWSEM_ET_DH, AWSEM_ET_BOND,
  AWSEM_N_ET
};

/* precision of the pair / switching arithmetic; force and energy
 * accumulation, reductions and the integrator are FP64 in both */
#define AWSEM_PREC_FP64 0
#define AWSEM_PREC_FP32 1

typedef struct awsem_ctx awsem_ctx;
typedef struct awsem_batch awsem_batch;

/* ---- single system: the FixBackbone drop-in ---------------------------- */

/* FixBackbone::FixBackbone(lmp, narg, arg) with arg[5] = coeff file and
 * arg[6] = sequence file (fb.cpp:86-1072).  Side files (anti_HB, anti_NHB,
 * para_HB, para_one, anti_one, ssweight, gamma.dat, burial_gamma.dat and
 * the fragment-memory .mem/.gro/gamma files) are read from workdir, which
 * plays the role of the LAMMPS working directory. */
int awsem_create(const char *coeff_path, const char *seq_path, const char *workdir,
                 int precision, awsem_ctx **out);

/* Same constructor with the per-system inputs given as arrays instead of
 * the sequence / ssweight / .mem / .gro files (coeff and the shared
 * parameter tables still come from files in param_dir).
 *   seq        : all chains concatenated (n letters), chain_len[nch]
 *   aps        : ssweight values [n][12] (row per residue, column per
 *                [SSWeight] flag; unflagged columns are ignored), or NULL
 *   structures : n_struct memory structures, the contents of .gro files:
 *                struct_len[s] residues, struct_xyz_nm[sum][2][3] the CA
 *                and CB coordinates in nm (multiplied by 10 like
 *                fm.cpp:87), struct_seq[sum] one-letter residue names
 *                ('G' residues carry no CB)
 *   memories   : n_mem lines of the [Memories] section (fb.cpp:1997-2008):
 *                structure index, target position, position in the
 *                structure (both 0-based), length, weight. */
int awsem_create_arrays(const char *coeff_path, const char *param_dir, int nch,
                        const int *chain_len, const char *seq, const double *aps,
                        int n_struct, const int *struct_len, const double *struct_xyz_nm,
                        const char *struct_seq, int n_mem, const int *mem_struct,
                        const int *mem_tpos, const int *mem_fpos, const int *mem_len,
                        const double *mem_weight, int precision, awsem_ctx **out);

/* Construct_Computational_Arrays (fb.cpp:1590-1698): per atom, the fix
 * group (0 = CA group, 1 = beta group, 2 = oxygen group, -1 none), the
 * molecule tag and the residue tag (both 1-based, atom_style awsemmd,
 * src/atom_vec_awsemmd.cpp:56). */
int awsem_set_topology(awsem_ctx *c, int n_atoms, const int *group, const int *mol,
                       const int *res);

/* FixBackbone::pre_force -> compute_backbone (fb.cpp:8181, 7047-7673) on
 * the device: f[n_atoms][3] += forces; energy[AWSEM_N_ET] receives the
 * per-term energies (compute_vector(k) == energy[k+1], fb.cpp:8223-8232).
 * When the excluded-volume pair style is enabled on the context
 * (awsem_set_vexcluded) its forces and energy (slot AWSEM_ET_VEXCLUDED)
 * are included, as PairExcludedVolume::compute (pev.cpp:50-138) would add
 * them in the same force evaluation. */
int awsem_compute(awsem_ctx *c, const double *x, double *f, double *energy);

/* Periodic box (LAMMPS domain->xperiodic/yperiodic/zperiodic and prd):
 * compute_backbone unwraps CA, CB and O with the atoms' image flags before
 * it evaluates anything (fb.cpp:7067-7112).  periodic[3] flags, prd[3] box
 * lengths (Angstrom, positive where periodic).  Default: not periodic. */
int awsem_set_box(awsem_ctx *c, const int *periodic, const double *prd);

/* awsem_compute with the atoms' LAMMPS image flags (atom->image, 10 bits
 * per dimension): a periodic coordinate is unwrapped as x + box * prd with
 * box = (image & 1023) - 512, (image >> 10 & 1023) - 512, (image >> 20) - 512
 * (fb.cpp:7067-7112); f += forces.  awsem_compute refuses a periodic box
 * (AWSEM_E_UNSUPPORTED), since it has no image flags to unwrap with. */
int awsem_compute_image(awsem_ctx *c, const double *x, const int32_t *image, double *f, double *energy);

/* energy.log of the reference (rank 0, fb.cpp:101-106 and 7661-7672): open
 * writes the tab-separated header; write appends "step " and then
 * "\t%8.6f" for the seventeen FixBackbone terms (EnergyTerms order,
 * fb.h:331-332) and their total, from energy[AWSEM_N_ET] or, when energy is
 * NULL, from the context's last evaluation.  The Excluded column is the
 * fix's own (zero in the reference's production build); the excluded-volume
 * pair style and the bonds are LAMMPS styles outside this log and its
 * total.  The caller decides the thermo steps (step % thermo_every == 0 in
 * the reference). */
int awsem_energy_log_open(awsem_ctx *c, const char *path);
int awsem_energy_log_write(awsem_ctx *c, int64_t step, const double *energy);
int awsem_energy_log_close(awsem_ctx *c);

/* Residue-layout variant of awsem_compute (no topology needed). */
int awsem_compute_res(awsem_ctx *c, const double *xres, double *fres, double *energy);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): awsem_compute_res with no allocation in the call.  The
 * per-call device buffers (coordinate staging, forces, energies and the
 * working slab) live in one fixed 4 MiB device array per GPU, held until the
 * process ends, so a system may have at most 9196 residues (larger ones get
 * AWSEM_E_NOMEM).  The context's constant device image must already exist:
 * call awsem_compute_res, awsem_compute or awsem_run once first (otherwise,
 * or after its parameters changed, AWSEM_E_STATE).  fres += forces as in
 * awsem_compute_res.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (other threads or contexts) race. */
int awsem_compute_res_noalloc(awsem_ctx *c, const double *xres, double *fres, double *energy);

/* pair_style vexcluded p rc_short rc_long + pair_coeff (pev.cpp:175-265):
 * tables are [6][6] indexed by atom type (1 CA, 2 N, 3 O, 4 CB, 5 H-beta).
 * enable=0 removes the pair style; NULL tables keep the deck defaults
 * (tools/create_project_tools/CoordinatesToWorkLammpsDataFile.py:314-322). */
int awsem_set_vexcluded(awsem_ctx *c, int enable, int p, const double *lambda,
                        const double *cut_short, const double *cut_long);

/* Harmonic bond style of the deck (K, r0 for CA-CA, CA-O, O-CA+1, CA-CB,
 * CA-Hbeta; CoordinatesToWorkLammpsDataFile.py:289-294); enable=0 removes it. */
int awsem_set_bonds(awsem_ctx *c, int enable, double k, const double *r0_5);

/* Neighbour-list skin of the terms that take their pairs from FixBackbone's
 * neighbour list with no cutoff of their own: Debye-Hueckel and the
 * solvent barrier (CA-CA pairs of the list, fb.cpp:8113-8117) and AMH-Go
 * (atom pairs of the full list, fb.cpp:3880-3895).  They count pairs closer
 * than pair_list_cutoff + skin, which is what a list LAMMPS has just built
 * with `neighbor skin bin` holds (the list request, fb.cpp:1805-1813).
 * Default 0 (a list rebuilt every step).  skin < 0: AWSEM_E_ARG. */
int awsem_set_list_skin(awsem_ctx *c, double skin);

/* Restrict accumulation to energy slots in mask (bit k = slot k); the
 * coeff file still decides which terms exist.  Test/debug aid. */
int awsem_set_term_mask(awsem_ctx *c, uint32_t mask);

/* Integer lists built on the device for one coordinate set (residue
 * layout).  kind: 0 density pairs of water well `arg`, 1 helix-well density
 * pairs, 2 DSSP (acceptor, donor) calls, 3 P_AP pairs, 4 helix (i,i+4)
 * pairs, 5 excluded-volume atom pairs, 6 water xi pairs of well `arg`.
 * Codes i*n+j (atom codes for kind 5), sorted ascending.  Returns the
 * count through *count; AWSEM_E_NOMEM if cap is too small. */
int awsem_list(awsem_ctx *c, int kind, int arg, const double *xres, int64_t *out,
               int cap, int *count);

/* FP64 velocity-Verlet + Langevin dynamics (LAMMPS fix nve + fix langevin
 * in real units; not part of the reference) with all force terms of the
 * context, kept resident on the device for nsteps.  t_start -> t_stop ramps
 * linearly over the call.  xres/vres are updated in place; fres holds the
 * force array as LAMMPS leaves it after the last step: every force term at
 * the final coordinates plus that step's fix langevin drag and random force
 * (the force the next step's first half-kick uses); energy the per-term
 * energies at the final coordinates. */
typedef struct awsem_langevin {
  double dt;        /* fs */
  double t_start;   /* K */
  double t_stop;    /* K */
  double damp;      /* fs */
  uint64_t seed;
  int64_t step0;    /* step counter of the first step (random stream) */
} awsem_langevin;

int awsem_run(awsem_ctx *c, int nsteps, const awsem_langevin *lg, double *xres,
              double *vres, double *fres, double *energy);

int awsem_n_residues(const awsem_ctx *c, int *n);
double awsem_pair_list_cutoff(const awsem_ctx *c);
int awsem_last_error(const awsem_ctx *c, char *buf, size_t n);
int awsem_destroy(awsem_ctx *c);

/* ---- batches of independent replicas ------------------------------------ */

/* A batch owns device copies of every replica's system and state.  The
 * contexts may be destroyed after the call.  All replicas must share one
 * coeff parameter set.  A context may be given more than once (replicas of
 * one system): its fragment-memory library, or its table curves, are then
 * stored on the device once. */
int awsem_batch_create(int n_ctx, awsem_ctx *const *ctxs, int precision,
                       awsem_batch **out);
int awsem_batch_size(const awsem_batch *b, int *n_replicas, int *n_residues_total);
/* host <-> device state, residue layout per replica concatenated in
 * replica order ([3][n_r][3] blocks); any pointer may be NULL.  f: after
 * awsem_batch_compute the forces of x; after awsem_batch_run as awsem_run's
 * fres (with the last step's Langevin drag and random force). */
int awsem_batch_set_state(awsem_batch *b, const double *x, const double *v);
int awsem_batch_get_state(awsem_batch *b, double *x, double *v, double *f,
                          double *energy /* [n_rep][AWSEM_N_ET] */);
/* one force evaluation of every replica (state x), results kept on device */
int awsem_batch_compute(awsem_batch *b);
/* nsteps of dynamics for every replica in one persistent launch
 * sequence; seed stream per replica = (seed, replica index). */
int awsem_batch_run(awsem_batch *b, int nsteps, const awsem_langevin *lg);
/* milliseconds of the last compute/run measured with HIP events on the
 * batch stream, and the main kernel's average duration */
int awsem_batch_timing(const awsem_batch *b, double *ms_total, double *ms_kernel_avg,
                       int *n_kernels);
/* Diagnostic: wall-clock ticks (100 MHz) accumulated per phase of the
 * resident engine when the process runs with AWSEM_PHASE_PROF=1:
 * [0..15] the first (longest) replica, [16..31] summed over replicas; slots
 * A, B (remainder), C, D, E, F, integrate-1, integrate-2, B fragment memory,
 * B local terms, B residue pairs, pair-list check/build, 4 unused.
 * AWSEM_E_STATE otherwise. */
int awsem_batch_phase_clock(const awsem_batch *b, uint64_t *ticks32);
/* Replays the resident launch from a captured HIP graph (default on; the
 * graph is recaptured when mode, dt or damp change). */
int awsem_batch_use_graph(awsem_batch *b, int on);
/* Verlet-list skin (Angstrom, default 2): lists hold residue pairs within
 * the largest term cutoff + extents + skin and are rebuilt when a residue
 * has moved by more than skin/2.  Any value gives the same forces. */
int awsem_batch_set_skin(awsem_batch *b, double skin);
/* Engine: 0 resident (one workgroup owns a replica for K steps, state in
 * LDS; default), 1 wide (every phase one launch over all replicas, state in
 * global memory; for few large replicas). */
int awsem_batch_set_engine(awsem_batch *b, int engine);
/* Autotuner: times probe_steps of dynamics (from the current state, which
 * is restored afterwards) for each engine, skin in {1, 1.5, 2, 2.5, 3} and
 * steps per launch in {5, 10, 20, 40, 80, 160}; keeps the fastest engine and skin and
 * reports them with the best steps per launch and its ms per batch-step. */
int awsem_batch_autotune(awsem_batch *b, const awsem_langevin *lg, int probe_steps,
                         double *skin, int *steps_per_launch, int *engine, double *ms_per_step);
/* Diagnostic: Verlet residue-pair lists of the resident engine summed over
 * replicas: list builds since creation, pairs in the current lists, and
 * replicas whose current list is invalid (overflowed or never built). */
int awsem_batch_list_stats(const awsem_batch *b, long long *builds, long long *pairs,
                           int *n_invalid);
/* Diagnostic (no counterpart in the reference, whose force loop is one
 * process per replica): fragment-memory sharing of the resident engine.
 * n_shared: replicas whose fragment-memory chunks idle workgroups may take
 * (0 when sharing is off: AWSEM_SHARE=0, a batch of one replica, or one too
 * large for LDS); chunks: chunks evaluated by another workgroup than the
 * replica's own since the batch was created. */
int awsem_batch_share_stats(const awsem_batch *b, int *n_shared, long long *chunks);
/* Diagnostic: the production pair selection of one replica as the last
 * evaluation of the batch used it (the same lists the force passes read):
 * kind 0 the Verlet list of residue pairs (a superset of every term's pairs,
 * i*n+j with i < j), 1 the DSSP calls (acceptor*n + donor; the pairs of
 * fb.cpp:8064-8094 that pass |O-N| < dssp_hdrgn_cut), 2 the P_AP pairs
 * (fb.cpp:8096-8103), 3 the residue pairs the excluded-volume pass visits,
 * 4 the residue pairs the water-xi and force passes visit (supersets: those
 * passes apply their exact cutoffs).  Sorted ascending; AWSEM_E_STATE when
 * the replica's list was not valid. */
int awsem_batch_lists(awsem_batch *b, int replica, int kind, int64_t *out, int cap, int *count);
int awsem_batch_last_error(const awsem_batch *b, char *buf, size_t n);
int awsem_batch_destroy(awsem_batch *b);

/* ---- serving: continuous batching of annealing jobs ---------------------
 * No counterpart in the reference (one LAMMPS run per process there).  Jobs
 * queue on the host; a scheduling round admits queued jobs in arrival order
 * while job slots and the device-memory budget allow, advances every active
 * job in one resident launch, and retires
 * finished jobs (final coordinates and per-term energies copied back, device
 * memory freed).  Each job runs fix nve + fix langevin with the target
 * temperature ramped from t_start to t_stop over its nsteps, with its own
 * random stream (seed), so its result does not depend on the other jobs in
 * the launches beyond floating-point accumulation order. */
typedef struct awsem_server awsem_server;
typedef struct {
  int precision;            /* AWSEM_PREC_FP32 or AWSEM_PREC_FP64 */
  int max_jobs;             /* job slots on the device */
  int steps_per_round;      /* steps per round of a job of median size; smaller
                               jobs run proportionally more (up to 8x), larger
                               ones fewer (at least 1) */
  double hbm_budget_bytes;  /* device memory for job images; <= 0: 80% of free */
  double dt, damp;          /* timestep and langevin damping (fs), all jobs */
} awsem_server_opts;
/* the force field (coeff + tables) of `forcefield` applies to every job */
int awsem_server_create(const awsem_ctx *forcefield, const awsem_server_opts *opts,
                        awsem_server **out);
/* queues a job: `system` gives sequence, memories and ssweight; xres/vres the
 * residue-layout start state (vres NULL = 0); returns its id */
int awsem_server_submit(awsem_server *s, const awsem_ctx *system, const double *xres,
                        const double *vres, int nsteps, double t_start, double t_stop,
                        uint64_t seed, uint64_t *job_id);
/* one scheduling round on the calling thread (not while started) */
int awsem_server_pump(awsem_server *s, int *n_active, int *n_queued);
/* a background thread runs rounds while jobs are queued or active */
int awsem_server_start(awsem_server *s);
int awsem_server_stop(awsem_server *s);
/* state: 0 queued, 1 running, 2 done, < 0 failed with that AWSEM_E_* code */
int awsem_server_poll(awsem_server *s, uint64_t job_id, int *state);
/* blocks until the job is done or failed (timeout_s <= 0: no limit) */
int awsem_server_wait(awsem_server *s, uint64_t job_id, double timeout_s);
/* final residue-layout coordinates, 19 energies and times_ms = {queued,
 * on device, total} in milliseconds */
int awsem_server_result(awsem_server *s, uint64_t job_id, double *xres, double *energy,
                        double *times_ms);
int awsem_server_release(awsem_server *s, uint64_t job_id);
/* stats[8]: rounds, jobs done, device bytes in use, peak active jobs,
 * replica-steps, device ms, budget bytes, active jobs */
int awsem_server_stats(awsem_server *s, double *stats);
int awsem_server_last_error(awsem_server *s, char *buf, size_t n);
int awsem_server_destroy(awsem_server *s);

/* ---- LAMMPS input decks (the commands the shipped decks use around
 * fix backbone) ---------------------------------------------------------- */

/* The subset of a LAMMPS input script that runs the FixBackbone path (the
 * shipped decks, e.g. examples/examples_new/1R69_single_memory/folding.in):
 * units real, dimension, timestep, boundary (non-periodic), neighbor /
 * neigh_modify / atom_modify (accepted: lists are the engine's own),
 * special_bonds fene, atom_style awsemmd, bond_style harmonic, pair_style
 * vexcluded + pair_coeff (pev.cpp:175-265), read_data (atom_style awsemmd:
 * id molecule residue type q x y z, src/atom_vec_awsemmd.cpp:56; Masses,
 * Atoms, Velocities, Bond Coeffs, Bonds, Pair Coeffs), mass, bond_coeff,
 * group (id / type / molecule lists and a:b ranges), velocity all create /
 * zero, fix nve, fix langevin, fix backbone (fb.cpp:86-106), fix print,
 * unfix, variable (equal / string / index; formulas with f_ID[k], emol,
 * epair, ke, pe, temp, step ...), thermo, thermo_style one / custom, dump
 * atom + dump_modify sort id, reset_timestep, print and run.  Anything
 * else is refused with AWSEM_E_UNSUPPORTED naming the command and line.
 * File names are relative to the script's directory, which is also the
 * FixBackbone working directory: energy.log (fb.cpp:101-106, 7661-7672)
 * and a thermo log, log.awsem, are written there.  The random streams of
 * velocity create and fix langevin are this library's, not LAMMPS's. */
typedef struct awsem_deck awsem_deck;
typedef struct awsem_deck_opts {
  int precision;       /* AWSEM_PREC_FP64 (0, default) or AWSEM_PREC_FP32 */
  int quiet;           /* no thermo / print lines on stdout */
  int64_t max_steps;   /* > 0: every run command runs at most this many steps */
} awsem_deck_opts;
/* Reads the script and executes it up to its first run command (no device
 * work).  On an error *out still holds the deck, for awsem_deck_last_error;
 * destroy it. */
int awsem_deck_open(const char *input_path, const awsem_deck_opts *opts, awsem_deck **out);
/* Executes the remaining commands: each run on the device (one replica of
 * a batch, continued across the output intervals), with its outputs. */
int awsem_deck_run(awsem_deck *d);
int awsem_deck_natoms(const awsem_deck *d, int *n_atoms, int *n_residues);
/* Atoms in ascending id order: ids[n], x[n][3], v[n][3]; the last
 * evaluated energies [AWSEM_N_ET]; the current step.  Any pointer may be
 * NULL. */
int awsem_deck_state(const awsem_deck *d, int64_t *ids, double *x, double *v, double *energy,
                     int64_t *step);
/* Atom id of every (kind, residue) slot of the residue layout ([3][n_res]:
 * CA, CB / H-beta, O), once fix backbone has been created. */
int awsem_deck_residue_atoms(const awsem_deck *d, int64_t *ids);
/* Value of an equal-style formula with the deck's current state. */
int awsem_deck_variable(awsem_deck *d, const char *formula, double *value);
int awsem_deck_last_error(const awsem_deck *d, char *buf, size_t n);
int awsem_deck_destroy(awsem_deck *d);

/* ---- the tertiary frustratometer ([Tertiary_Frustratometer]) ------------
 * compute_tert_frust / compute_tert_frust_singleresidue (fb.cpp:5057-5412)
 * at one coordinate set (residue layout), with the coeff section's cutoff,
 * decoy count and mode.  Decoys draw from the context's own copy of the C
 * library's unseeded rand() stream (glibc's generator, seed 1), continued
 * across calls; the configurational decoy statistics are computed once, at
 * the first call with a contact (fb.cpp:5099, 5342).  rows: 11 doubles per
 * contact (i, j, i_chain, j_chain, r_ij, rho_i, rho_j, native energy, decoy
 * mean, decoy std, frustration index; 1-based) or per residue in
 * singleresidue mode (j, j_chain, r_ij and rho_j 0); *n_rows receives the
 * count (AWSEM_E_ARG when it exceeds cap, the rand() stream then left where
 * it was). */
int awsem_tert_frust(awsem_ctx *c, const double *xres, double *rows, int cap, int *n_rows);
/* The reference's output when step % output interval == 0:
 * tertiary_frustration.dat and .tcl in the context's directory (created
 * with their header at the first call), "# timestep: step" and the rows in
 * its formats (fb.cpp:855-867, 7586-7596, 5104-5124, 5168-5182). */
int awsem_tert_frust_write(awsem_ctx *c, const double *xres, int64_t step);
/* The output interval of the [Tertiary_Frustratometer] section (its third
 * line, fb.cpp:385-397), 0 without the section: a driver stops at its
 * multiples to call awsem_tert_frust_write. */
int awsem_tert_frust_every(const awsem_ctx *c, int64_t *every);

/* The n-mer frustratometer ([Nmer_Frustratometer], fb.cpp:398-412,
 * 5600-6061) at xres, on the same rand() stream.  rows: 7 doubles per row
 * (pairwise: i, j, contacts, native energy, decoy mean, decoy std,
 * frustration index, for the n-mer pairs with enough contacts; singlenmer:
 * i, 0, 0, ... for every n-mer; 1-based starts).  With the trap flag
 * (pairwise), traps: 7 doubles per trap (row, first n-mer, second n-mer, k,
 * backward, threshold energy, trap energy) in the reference's order.
 * *n_rows / *n_traps receive the counts; AWSEM_E_ARG when either exceeds
 * its cap, the rand() stream then left where it was. */
int awsem_nmer_frust(awsem_ctx *c, const double *xres, double *rows, int cap, int *n_rows, double *traps, int tcap,
                     int *n_traps);
/* The reference's output when step % output interval == 0:
 * nmer_frustration.dat, nmer_frustration.tcl and (trap flag) nmer_traps.dat
 * in the context's directory (fb.cpp:885-896, 7592-7603, 5626-5706,
 * 5855-5890). */
int awsem_nmer_frust_write(awsem_ctx *c, const double *xres, int64_t step);
/* The output interval of the [Nmer_Frustratometer] section, 0 without it. */
int awsem_nmer_frust_every(const awsem_ctx *c, int64_t *every);

/* The fragment frustratometer ([Fragment_Frustratometer], fb.cpp:359-384,
 * 4431-4832) at xres.  Shuffle mode: the native memory energies and
 * n_calcs - 1 placements of the decoy fragments drawn from the rand()
 * stream; read mode: the native memory energies against decoy-structure
 * statistics computed at the first call (at step 0 in the reference).
 * rows: n x 3 (frustration index (native - mean) / std, native - mean, std
 * per residue).  gen (may be NULL; n * (n_decoys + 1) doubles): at the call
 * that computes read mode's statistics, the n native energies then the
 * n x n_decoys decoy energies (residue-major), and *gen_written = 1. */
int awsem_frag_frust(awsem_ctx *c, const double *xres, double *rows, double *gen, int *gen_written);
/* The reference's output when step % output interval == 0:
 * fragment_frustration.dat, _gap.dat, _variance.dat (one line per call),
 * _native.dat and _decoy.dat (read mode's generated energies) in the
 * context's directory (fb.cpp:794-800, 4676-4688, 4786-4802). */
int awsem_frag_frust_write(awsem_ctx *c, const double *xres, int64_t step);
/* mode (0 shuffle, 1 read, -1 without the section), decoy calculations,
 * decoy fragments or structures, output interval (0 without the section). */
int awsem_frag_frust_info(const awsem_ctx *c, int *mode, int *n_calcs, int *n_decoys, int64_t *every);

/* ---- order parameters (compute qwolynes, compute qonuchic) ---------------
 * Replace `compute ID alpha_carbons qwolynes rnative.dat sep exp`
 * (src/compute_q_wolynes.cpp:51-88, compute_scalar 171-236) and
 * `compute ID alpha_carbons qonuchic {cutoff rc native.dat factor |
 * cutoff/gauss rc native.dat | shadow map factor | shadow/gauss map}`
 * (src/compute_q_onuchic.cpp:35-93, 150-317) for the n CAs of a system in
 * residue order.  Evaluated on the device, for one trace or for replicas of
 * a batch at their current positions.  On an error *out still holds the
 * object, for awsem_q_last_error; destroy it. */
#define AWSEM_Q_WOLYNES 0
#define AWSEM_Q_ONUCHIC_CUTOFF 1
#define AWSEM_Q_ONUCHIC_CUTOFF_GAUSS 2
#define AWSEM_Q_ONUCHIC_SHADOW 3
#define AWSEM_Q_ONUCHIC_SHADOW_GAUSS 4
typedef struct awsem_q awsem_q;
int awsem_q_wolynes_create(const char *rnative_path, int n, int sep, double sigma_exp, awsem_q **out);
/* kind AWSEM_Q_ONUCHIC_*; path: native CA coordinates ("n" then n rows
 * of x y z) for the cutoff kinds, the contact map ("i j rn" rows, 1-based)
 * for the shadow kinds; r_contact for the cutoff kinds, factor for cutoff
 * and shadow */
int awsem_q_onuchic_create(int kind, const char *path, int n, double r_contact, double factor, awsem_q **out);
/* Q of one CA trace xca[n][3] (the coordinates as the compute reads them:
 * compute qwolynes uses the stored, wrapped ones, its unwrap being
 * overwritten at compute_q_wolynes.cpp:198-214) */
int awsem_q_compute(awsem_q *q, const double *xca, double *value);
/* Q of replicas [first, first + count) of a batch (each with the native's
 * residue count) at their current device positions: values[count] */
int awsem_batch_q(awsem_batch *b, awsem_q *q, int first, int count, double *values);
int awsem_q_last_error(const awsem_q *q, char *buf, size_t n);
int awsem_q_destroy(awsem_q *q);

#ifdef __cplusplus
}
#endif
#endif
