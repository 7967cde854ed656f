/*
 * awsem_oracle.h -- CPU FP64 restatement of AWSEM-MD's FixBackbone force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in awsemmd_amd/.  Only tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Every routine follows the reference loops it cites (paths relative to the
 * reference repository root):
 *   src/fix_backbone.cpp   (fb.cpp)   FixBackbone ctor, compute_backbone, compute_pair
 *   src/smart_matrix_lib.h (sml.h)    cWell / cP_AP / cR switching functions
 *   src/fragment_memory.cpp (fm.cpp)  Fragment_Memory, Gamma_Array
 *   src/pair_excluded_volume.cpp      PairExcludedVolume::compute
 * The harmonic bonds and the Langevin integrator live in LAMMPS core (not in
 * the reference); they are restated from LAMMPS' published formulas and are
 * "parity unpinned" against the reference (pinned only by closed forms,
 * finite differences and statistics).
 *
 * Single-rank semantics: every residue is LOCAL, ghost atoms do not exist,
 * residue k of the sequence file is residue index k (0-based) and its
 * chain is the sequence-file line it belongs to.
 */
#ifndef AWSEM_ORACLE_H
#define AWSEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Energy slots: same order as FixBackbone::EnergyTerms (fb.h:331-332),
 * plus one extra slot for the harmonic bond style (outside the fix). */
enum {
  OR_ET_TOTAL = 0, OR_ET_CHAIN, OR_ET_SHAKE, OR_ET_CHI, OR_ET_RAMA,
  OR_ET_VEXCLUDED, OR_ET_DSSP, OR_ET_PAP, OR_ET_WATER, OR_ET_BURIAL,
  OR_ET_HELIX, OR_ET_AMHGO, OR_ET_FRAGMEM, OR_ET_VFRAGMEM, OR_ET_CONT_REST,
  OR_ET_MEMB, OR_ET_SSB, OR_ET_DH, OR_ET_BOND, OR_N_ET
};

typedef struct or_ctx or_ctx;

/* Branch counters: how often each selection branch of the DSSP, P_AP and
 * helix terms ran since the last reset (evidence that a test system
 * exercises them).  DSSP: pairs evaluated per class 1..4 (fb.cpp:2956-3002),
 * class 2 with the beta prediction, repulsive / antiparallel / parallel
 * channels on with theta > 1e-3, nu switches of i and j evaluated, pairs
 * leaving at nu < 1e-4 (fb.cpp:3052); P_AP: medium, long (same chain, other
 * chain), parallel (same chain, other chain) bands and beta-predicted pairs
 * (fb.cpp:3295-3315); helix pairs evaluated (fb.cpp:3706). */
enum {
  OR_BR_DSSP_CLASS1 = 0, OR_BR_DSSP_CLASS2, OR_BR_DSSP_CLASS3, OR_BR_DSSP_CLASS4, OR_BR_DSSP_BETA,
  OR_BR_DSSP_REP, OR_BR_DSSP_AP, OR_BR_DSSP_P, OR_BR_DSSP_NU0, OR_BR_DSSP_NU1, OR_BR_DSSP_NU_EXIT,
  OR_BR_PAP_MED, OR_BR_PAP_LONG, OR_BR_PAP_LONG_X, OR_BR_PAP_PAR, OR_BR_PAP_PAR_X, OR_BR_PAP_BETA,
  OR_BR_HELIX, OR_N_BRANCH
};
int or_branch_counts(or_ctx *c, long long *out, int n, int reset);

/* Reads the coeff file, the sequence file and the fixed-name side files
 * (anti_HB, anti_NHB, para_HB, para_one, anti_one, ssweight, gamma.dat,
 * burial_gamma.dat, fragment memory files) from workdir, like the
 * FixBackbone constructor (fb.cpp:86-1072).  Returns 0 or a negative code;
 * err receives the reference's message. */
int or_create(const char *coeff_path, const char *seq_path, const char *workdir,
              or_ctx **out, char *err, int errlen);
void or_destroy(or_ctx *c);

int or_n_residues(const or_ctx *c);
/* chain index (0-based) of residue i, and sequence letter */
int or_chain_of(const or_ctx *c, int i);
char or_seq_at(const or_ctx *c, int i);
/* pair_list_cutoff of fb.cpp:1048-1066 */
double or_pair_list_cutoff(const or_ctx *c);

/* Neighbour-list skin of the terms that take their pairs from FixBackbone's
 * neighbour list without a cutoff of their own (Debye-Hueckel, the CA-CA
 * pairs of the solvent barrier, AMH-Go): pairs closer than
 * pair_list_cutoff + skin count, as in a list LAMMPS has just built
 * (fb.cpp:1805-1813).  Default 0. */
void or_set_list_skin(or_ctx *c, double skin);

/* Excluded-volume pair style settings (pair_style vexcluded p rs rl +
 * pair_coeff per type pair, pev.cpp:175-265).  Atom types: 1 CA, 2 N,
 * 3 O, 4 CB, 5 H-beta (Gly).  Defaults are the deck written by
 * CoordinatesToWorkLammpsDataFile.py:314-322. */
void or_set_vexcluded(or_ctx *c, int p, const double *lambda6x6,
                      const double *cut_short6x6, const double *cut_long6x6);

/* Harmonic bond style of a deck (bond_style harmonic + bond_coeff): one K
 * and r0 for CA-CA, CA-O, O-CA(+1), CA-CB, CA-H-beta.  Defaults: the deck
 * written by CoordinatesToWorkLammpsDataFile.py:289-294 (K 200). */
void or_set_bonds(or_ctx *c, double k, const double *r0_5);

/* Forces are ADDED into f* (pre_force semantics).  energy[OR_N_ET] is
 * overwritten.  term_mask selects energy slots (bit k = slot k) whose
 * forces and energies are accumulated; the coeff flags still decide which
 * terms exist.  Bit OR_ET_VEXCLUDED adds PairExcludedVolume, bit OR_ET_BOND
 * adds the harmonic bonds of the deck. */
int or_compute(or_ctx *c, const double *xca, const double *xcb, const double *xo,
               double *fca, double *fcb, double *fo, double *energy,
               unsigned term_mask);

/* Integer lists (sorted, pairs as i*n+j codes written to out, returns the
 * count, or -needed if cap is too small).
 *  kind 0: density pairs i<j for water well w (arg) (fb.cpp:7765)
 *  kind 1: density pairs i<j for the helix well (fb.cpp:7775)
 *  kind 2: DSSP calls (acceptor i, donor j) that pass rNO<=cut (fb.cpp:8064-8094, 2904)
 *  kind 3: P_AP calls (i<j) with rsq<pap_cutoff_sq (fb.cpp:8098)
 *  kind 4: helix pairs (i, i+4) that reach compute_helix_dtheta_pair's body
 *  kind 5: excluded-volume atom pairs within cutoff (atom code = 3*res+kind)
 *  kind 6: water DL2 xi pairs for well w (fb.cpp:7925 predicate) */
int or_list(or_ctx *c, int kind, int arg, const double *xca, const double *xcb,
            const double *xo, long long *out, int cap);

/* Fragment-memory introspection (fm.cpp) */
int or_fm_count(const or_ctx *c);
/* fragment-memory table mode: last error of or_compute (an out-of-range
   distance), grid size / neighbour count, one curve's energy and force */
const char *or_last_error(const or_ctx *c);
void or_fm_table_dims(const or_ctx *c, int *tb_size, int *tb_nbrs);
int or_fm_table(const or_ctx *c, int itb, double *e, double *f);
int or_fm_info(const or_ctx *c, int k, int *pos, int *len, double *weight);
double or_fm_rf(const or_ctx *c, int k, int ires, int iatom, int jres, int jatom);
char or_fm_se(const or_ctx *c, int k, int resno);
double or_fm_gamma(const or_ctx *c, int ires_type, int jres_type, int ires, int jres);
int or_fm_minsep(const or_ctx *c);
int or_fm_maxsep(const or_ctx *c);
int or_fm_map(const or_ctx *c, int i, int *out, int cap);

/* Switching functions exposed for pinning against the reference headers */
void or_theta_pair(const or_ctx *c, int helix, int well, double r,
                   double *theta, double *prd_theta);
void or_pap_nu(const or_ctx *c, double drsq, double *nu, double *prd_nu);

/* LAMMPS-style fix nve + fix langevin step in real units (FP64), driving
 * or_compute with every force term of the context.  x/v are n*3*3 arrays
 * laid out [kind][residue][xyz] with kind 0 CA, 1 CB, 2 O.  f holds the
 * force of the previous step on entry and of this step on exit.  Random
 * numbers come from the same counter hash as the GPU build (seed, replica,
 * step, atom). */
int or_langevin_run(or_ctx *c, int nsteps, double dt, double t_start,
                    double t_stop, double damp, unsigned long long seed,
                    int replica, long long step0, double *x, double *v,
                    double *f, double *energy);

/* Order parameters of a CA trace (xca: n x 3): compute qwolynes
 * (src/compute_q_wolynes.cpp) against n x n native distances, and compute
 * qonuchic (src/compute_q_onuchic.cpp): type 0 cutoff / 1 cutoff/gauss
 * from native coordinates nat (n x 3) and contact radius rc, type 2 shadow
 * / 3 shadow/gauss from a contact map (0-based i, j, native distance). */
double or_q_wolynes(int n, const double *xca, const double *rn, int sep, double sexp);

/* The tertiary frustratometer ([Tertiary_Frustratometer], fb.cpp:5057-5412)
 * at one coordinate set: rows of 11 doubles (i, j, i_chain, j_chain, r_ij,
 * rho_i, rho_j, native, decoy mean, decoy std, frustration index; j = 0 in
 * singleresidue mode); returns the rows or -needed.  The decoys draw from
 * the context's rand() stream (glibc's, unseeded: or_rng_seed(c, 1) at
 * creation), which continues across calls, as do the configurational decoy
 * statistics (computed once, fb.cpp:5099, 5342). */
int or_tert_frust(or_ctx *c, const double *xca, const double *xcb, const double *xo, double *rows, int cap);
/* the n-mer frustratometer ([Nmer_Frustratometer], fb.cpp:5600-6061) on
 * the same rand() stream: rows of 7 doubles (i, j, contacts, native, decoy
 * mean, decoy std, index; singlenmer: i, 0, 0, ...) and, with the trap
 * flag, trap records of 7 doubles (row, first, second, k, backward,
 * threshold, energy); returns the rows (-needed beyond cap) */
int or_nmer_frust(or_ctx *c, const double *xca, const double *xcb, const double *xo, double *rows, int cap,
                  double *traps, int tcap, int *ntr);
/* the fragment frustratometer ([Fragment_Frustratometer], fb.cpp:4431-4832)
 * on the same rand() stream: rows n x 3 (index, native - mean, std per
 * residue); read mode's generated energies (n natives, n x n_decoys) into
 * gen at the call that computes them (*gen_written 1); < 0 on errors */
int or_frag_frust(or_ctx *c, const double *xca, const double *xcb, const double *xo, long long step, double *rows,
                  double *gen, int *gen_written);
int or_frag_frust_dims(const or_ctx *c, int *mode, int *ncalc, int *ndecoys);
void or_rng_seed(or_ctx *c, unsigned seed);
int or_rng_next(or_ctx *c);
double or_q_onuchic(int n, const double *xca, int type, const double *nat, double rc, int nmap, const int *mi,
                    const int *mj, const double *mrn, double factor);

#ifdef __cplusplus
}
#endif
#endif
