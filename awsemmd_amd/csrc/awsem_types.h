// awsem_types.h -- plain-data structures shared by the host code and the
// gfx950 kernels of the FixBackbone force path.
#pragma once
#include <stdint.h>

#define AW_N_ET 19
#define AW_MAXCH 1000

// energy slots (FixBackbone::EnergyTerms, src/fix_backbone.h:331-332) + bonds
enum {
  AW_ET_TOTAL = 0, AW_ET_CHAIN, AW_ET_SHAKE, AW_ET_CHI, AW_ET_RAMA, AW_ET_VEXCLUDED,
  AW_ET_DSSP, AW_ET_PAP, AW_ET_WATER, AW_ET_BURIAL, AW_ET_HELIX, AW_ET_AMHGO,
  AW_ET_FRAGMEM, AW_ET_VFRAGMEM, AW_ET_CONT_REST, AW_ET_MEMB, AW_ET_SSB, AW_ET_DH,
  AW_ET_BOND
};

// per-residue info word
#define RI_TYPE_MASK 31
#define RI_FIRST 32
#define RI_LAST 64
#define RI_GLY 128
#define RI_PRO 256
#define RI_PHOS 512   // phosphorylated ([Phosphorylation], fb.cpp:643-682)
#define RI_CHAIN_SHIFT 16

// Force-field parameters after the constructor's parsing and epsilon
// scaling (src/fix_backbone.cpp:165-516, src/smart_matrix_lib.h:363-368,
// 111-115).  Lives in device global memory; kernels read it through a
// pointer (it is ~55 KB, too large for kernel arguments).
struct FFParams {
  int chain_flag, shake_flag, chi_flag, rama_flag, rama_p_flag, ssweight_flag;
  int dssp_flag, pap_flag, water_flag, burial_flag, helix_flag, fm_flag;
  double epsilon;
  double an, bn, cn, ap, bp, cp, ah, bh, ch;
  double k_chain[3], r_ncb0, r_cpcb0, r_ncp0;
  double k_shake, r_sh1, r_sh2, r_sh3;
  double k_chi, chi0, k_rama, k_water;
  int n_rama_par, n_rama_p_par;
  int ssweight[12];
  double w[12], phiw[12], phi0[12], psiw[12], psi0[12];
  double cphi0[12], sphi0[12], cpsi0[12], spsi0[12];  // cos/sin of phi0, psi0
  // DSSP hydrogen bonds
  double k_dssp, hbscl[4][9], NO_zero, HO_zero, sigma_NO_sqinv, sigma_HO_sqinv;
  double dssp_cut, dssp_cut_sq, pref[2], d_nu0, nu_cut1_sq, nu_cut2_sq;
  double anti_HB[20][20][2], anti_NHB[20][20][2], para_HB[20][20][2];
  double para_one[20], anti_one[20];
  // P_AP
  double k_global_pap, k_betapred_pap, k_pap[3], pap_cut, pap_pref, pap_cutoff_sq;
  double pap_drmax_sq, pap_drmin_sq;
  int i_med_min, i_med_max, i_diff_pap;
  // water
  int contact_cutoff, n_wells, well_flag[5];
  double water_kappa, water_kappa_sigma, treshold;
  double well_r_min[5], well_r_max[5], w_rmin_th[5], w_rmax_th[5], w_rmin_th_sq[5], w_rmax_th_sq[5];
  double water_gamma[5][20][20][2];
  // burial
  double k_burial, burial_kappa, burial_ro_min[3], burial_ro_max[3], burial_gamma[20][3];
  // helix
  int helix_i_diff, pro_accepter_flag;
  double k_helix, helix_gamma_p, helix_gamma_w, helix_kappa, helix_kappa_sigma, helix_treshold;
  double helix_cutoff, helix_cutoff_sq, helix_r_min, helix_r_max;
  double h_rmin_th, h_rmax_th, h_rmin_th_sq, h_rmax_th_sq;
  double h4prob[20], h4prob_pro_accepter;
  double helix_NO_zero, helix_HO_zero, helix_sigma_NO_sqinv, helix_sigma_HO_sqinv;
  // fragment memory
  double k_frag_mem, fm_sigma_exp;
  int fm_min_sep, fm_max_sep;
  // fragment-memory table mode ([Fragment_Memory_Table], fb.cpp:310-320):
  // curves on the grid r = tb_rmin + ir tb_dr, ir < tb_size
  int fm_tb_flag, fm_use_table, tb_size, pad_tb;
  double tb_rmin, tb_rmax, tb_dr, tb_well_width;
  // excluded volume pair style (src/pair_excluded_volume.cpp)
  int ev_on, ev_p;
  double ev_lambda[6][6], ev_cut_short[6][6], ev_cut_long[6][6];
  // harmonic bonds of the deck
  int bond_on;
  double bond_k, bond_r0[5];
  // coeff-toggled terms (fb.cpp:297-358, 412-421, 466-478), strengths
  // after the epsilon scaling of fb.cpp:502-516 (AMH-Go, contact
  // restraints); the others carry epsilon where they are evaluated
  int amhgo_flag, cr_flag, vfm_flag, ssb_flag, memb_flag, phosph_flag, dh_flag, xt_on;
  double k_amh_go, amh_go_p, amh_go_rc;
  int amh_go_censor, pad_go;
  double k_cont_rest, cr_sigma, cr_sigma_sq_inv, cr_dr_cutoff, cr_glob_cutoff_sq;
  double k_vfm, vfm_sigma, vfm_sigma_sq;
  double k_ssb1, ssb_rmin1, ssb_rmax1, k_ssb2, ssb_rmin2, ssb_rmax2, ssb_kappa, ssb_rshift[20];
  int ssb_ij_sep, ssb_rad_cor;
  double k_memb, memb_kbin, memb_xo[3], memb_len, memb_rho0_max, memb_rho0_distor, g_memb[3][4];
  int memb_pore_type, n_phosph;
  double k_hypercharge;
  int phosph_res[20];
  double dh_k[3];            // k++, k--, k+- (fb.cpp:468)
  double dh_kscreen, dh_length;
  int dh_min_sep, pad_dh;
  // the hypercharged water table of phosphorylated pairs (fb.cpp:643-669)
  double phosph_gamma[5][20][20][2];
  // [Tertiary_Frustratometer] (fb.cpp:385-397): contact cutoff, decoys,
  // output interval, mode 0 configurational / 1 mutational / 2 single residue
  int tf_flag, tf_ndecoys, tf_freq, tf_mode;
  double tf_cutoff;
  // [Nmer_Frustratometer] (fb.cpp:398-412): n-mer size, contact cutoff,
  // contacts needed, decoys, output interval, tcl thresholds and neutral
  // flag, trap flags, trap width in sigmas, secondary-structure fraction,
  // mode 0 pairwise / 1 singlenmer
  int nf_flag, nf_size, nf_contacts, nf_ndecoys, nf_freq, nf_neutral, nf_trap, nf_draw_trap, nf_mode, pad_nf;
  double nf_cutoff, nf_min_thr, nf_high_thr, nf_trap_sigma, nf_ss_frac;
  // [Fragment_Frustratometer] (fb.cpp:359-384): mode 0 shuffle / 1 read,
  // decoy calculations (shuffle; read: decoys + 1), output interval, the
  // sequence-separation scaling flag and exponent and the well width (read
  // mode; shuffle mode leaves them unset in the reference: no scaling,
  // width 1 here), the normalisation flag (no effect in the reference)
  int ff_flag, ff_mode, ff_ncalc, ff_freq, ff_seqsep, ff_norm;
  double ff_ww, ff_seqsep_gamma;
  // terms whose pairs come from FixBackbone's neighbour list with no cutoff
  // of their own (Debye-Hueckel, the barrier's CA-CA pairs, AMH-Go) count
  // pairs closer than pair_list_cutoff + list_skin (a list LAMMPS has just
  // built, fb.cpp:1805-1813); list_r2 is that distance squared
  double list_skin, list_r2;
  // derived cutoffs
  double pair_list_cutoff;   // src/fix_backbone.cpp:1048-1066
  double nl_skin;            // Verlet-list skin of the resident engine (A)
  double pair_reach;         // largest pair cutoff of any term (density, DSSP, P_AP, EV)
  double dens_reach;         // largest density-well outer cutoff (water/helix)
  double ev_reach;           // largest excluded-volume cutoff
  uint32_t term_mask;
};

// Fragment memory work item: one (i, j) residue pair of one group of
// memories sharing a window (pos, len); every memory of the group covers
// the same pairs.  An item carries its residues, the index of its Rf entry
// for the group's first memory (float4/double4 units; the entry of memory m
// is rf0 + m * stride) and the group's memory weights (coef0: start of a
// zero-padded run of AW_FM_PAD weights or more), so a work item needs one
// dependent load before its Rf loads.  The Gaussian width enters through
// sc = sqrt(log2(e) / (2 sigma^2)): the device Rf entries of an item are
// stored pre-multiplied by sc, the item scales its four distances by sc, and
// exp(-dr^2 / (2 sigma^2)) = exp2(-(sc r - sc rf)^2).  An entry whose
// Gaussian is skipped (a glycine CB of the memory or of the target pair,
// fb.cpp:4148-4149) holds AW_FM_FAR, whose Gaussian underflows to exactly 0.
// invs = 1 / (sigma^2 sc) turns the sum of w e (sc r - sc rf) into the
// force factor.
#define AW_FM_PAD 16
#define AW_FM_FAR 1.0e6
#define AW_FM_GLY_I (1 << 30)  // flags in nmem: residue i / j is glycine
#define AW_FM_GLY_J (1 << 29)
#define AW_FM_LOGW (1 << 28)   // flag in nmem: the device weights of the group are log2(w), all w > 0
#define AW_FM_NMEM_MASK ((1 << 28) - 1)
struct FMItem {
  float gamma, sc, invs;     // gamma, sqrt(log2(e)/(2 sigma^2)), 1/(sigma^2 sc)
  int ij;                    // (i << 16) | j, replica-local residue indices
  int rf0, stride, nmem, coef0;
};
struct FMItemD {
  double gamma, sc, invs;
  int ij;
  int rf0, stride, nmem, coef0, pad;
};

// One replica (system) as the kernels see it: device pointers to its own
// data and its dynamics schedule for the current launch.
// Verlet list block: [0] count, [1] valid, [2] builds, [3] -, [4..11] list
// and queue counters when kept in global memory, then the list (nl_cap
// packed pairs) and the work queues (5 nl_cap)
#define AW_NL_HDR 16
// systems above this size get a cell-list table for their list builds
// (AWSEM_CELL_MIN_N overrides it).  Below it the all-pairs build wins: its
// pair order keeps neighbouring lanes on neighbouring residues.
#define AW_CELL_MIN_N 4096
// Fragment-memory sharing (resident engine): the fragment-memory items of a
// large replica are cut into chunks of consecutive items, each touching
// residues [lo, lo + AW_SHARE_SPAN) only.  While the replica's workgroup
// runs its other passes, workgroups that are done with their own replicas
// take chunks, evaluate them against a copy of the replica's CA and CB
// positions and add the forces into the replica's accumulator; the owner
// takes what is left, waits for the chunks taken elsewhere and adds the
// accumulator into its forces before it integrates.
#define AW_SHARE_SPAN 64
#define AW_SHARE_MAXCH 64      // chunks per replica (one 64-bit claim mask)
#define AW_SHARE_TAKE 8        // chunks one helping workgroup takes at once (one per wave)
struct ShareChunk {
  int k0, k1;                  // items [k0, k1)
  int lo, pad;                 // first residue of the span
};
#define AW_RF_F_VALID 1   // f holds the forces of x: dynamics continue from them
#define AW_RF_NL_RESET 2  // x was replaced: rebuild the Verlet list first
struct ReplicaDesc {
  int n, nch;             // residues, chains
  int fm_items, nl_cap;   // fragment-memory items; Verlet list capacity (pairs)
  const int *rinfo;       // [n]
  const double *aps;      // [n][12]
  const int *ch_start, *ch_len;  // [nch]
  double *x, *v, *f;      // [3][n][3] each (CA, CB, O)
  double *energy;         // [AW_N_ET]
  double *scratch;        // working slab in global memory (phase engine; replicas too large for LDS)
  const void *items, *rf, *coef;  // FMItem/FMItemD [fm_items], float4/double4 Rf, weights
  int *nl;                // Verlet list block: AW_NL_HDR header ints, nl_cap pairs, 5 nl_cap queue slots
  double *nlx;            // [4 n]: CA positions and extents at the last list build
  int *cells;             // cell-list table for the list build (wide engine): counts [cell_h],
                          // offsets [cell_h], members [n], bucket [n], slot [n], packed cell [n];
                          // null: all-pairs build
  int cell_h, pad_c;      // buckets (a power of two)
  // dynamics of this replica in the current launch
  double t_start, t_stop;      // target temperature, ramped linearly over the run
  unsigned long long seed;     // random stream (seed, rng_id, step, atom, component)
  long long step0;             // stream step index of the launch's first step
  int ramp0, run_steps;        // steps of the run before this launch; length of the run
  int nsteps;                  // steps of this replica in this launch
  int rng_id;
  int flags, pad;              // AW_RF_*
  // fragment-memory sharing (sh_n chunks; 0: the replica evaluates its own)
  unsigned long long *sh_word; // (energies wanted << 32) | next unclaimed chunk
  int *sh_done;                // chunks finished elsewhere in the current evaluation
  const ShareChunk *sh_chunks; // [sh_n]
  double *sh_x;                // [2][n][3]: CA and CB positions of the current evaluation
  double *sh_acc;              // [2][n][3] forces of the chunks finished elsewhere, then their energy
  int sh_n, sh_q;              // chunks; position in the batch's list of sharing replicas
  // fragment-memory table mode: one item of 4 ints per evaluated atom pair
  // (i, j, atom pair 0..3, curve id or -1: range check only), curves of
  // tb_size (energy, force) pairs in the pair arithmetic's type
  int tb_items, pad_tb;
  const int *tb_item;
  const void *tb_curve;
  int *fault;                  // device error word: bit 0 an FM table distance outside [rmin, rmax]
  // coeff-toggled terms (null / 0 when off): AMH-Go items of each atom
  // a = 2 i + (0 CA, 1 CB) at go_start[a] .. go_start[a + 1); contact
  // restraints, Debye-Hueckel pairs and vector fragment-memory items;
  // membrane classes per residue
  int n_cr, n_dh, n_vfm, n_go;
  const int *go_start;
  const void *go_item, *cr_item, *dh_item, *vfm_item;
  const int *zres;
  double go_norm;
  // the system's list reach of the list-limited terms (pair_list_cutoff +
  // list skin, squared) and the contact restraints' global cutoff squared:
  // both follow the system's own files
  double list_r2, cr_cut2;
};

// order-parameter pair (compute qwolynes / qonuchic): gauss kinds add
// exp(-(r - a)^2 b), cutoff kinds count r^2 < a
struct QItem {
  int i, j;
  double a, b;
};

// coeff-toggled term items (FP64: these terms are few per residue)
struct XtGo {                  // one directed AMH-Go pair a -> b (fb.cpp:3895-3950)
  int b, pad;                  // atom 2 j + (0 CA, 1 CB)
  double rn, gamma, s2;        // native distance, gamma, sigma^2 = |i - j|^0.3
};
struct XtPair {                // contact restraint (w, r0) / Debye-Hueckel (k q_i q_j, -) / vector FM (gf, -)
  int i, j;
  double a, b;
};
#define AW_FAULT_TB_RANGE 1
