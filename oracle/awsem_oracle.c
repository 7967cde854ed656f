/*
 * awsem_oracle.c -- CPU FP64 restatement of AWSEM-MD's FixBackbone path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see
 * awsem_oracle.h.  Function headers cite the reference file:line each one
 * restates; fb.cpp = src/fix_backbone.cpp, sml.h = src/smart_matrix_lib.h,
 * fm.cpp = src/fragment_memory.cpp, pev.cpp = src/pair_excluded_volume.cpp.
 *
 * Single-rank, non-periodic semantics: all residues LOCAL, no ghosts, the
 * LAMMPS half neighbour list is replaced by exhaustive enumeration of every
 * unordered atom pair (the reference's predicates select the same pairs
 * because every cutoff used is <= pair_list_cutoff).
 */
#include "awsem_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>

#define MAXCH 1000
#define DELTA 0.00001          /* fb.cpp:40 */
#define DELTA_WATER_XI 1e-8    /* fb.cpp:41 */
#define DELTA_HELIX_XI 1e-6    /* fb.cpp:42 */
#define DSSP_NU_DELTA 1e-4     /* fb.cpp:43 */
#define PAP_DELTA 1e-12        /* fb.cpp:44 */

/* fb.cpp:54 residue letter -> type index */
static const int se_map[26] = {0, 0, 4, 3, 6, 13, 7, 8, 9, 0, 11, 10, 12, 2, 0, 14, 5, 1, 15, 16, 0, 19, 17, 0, 18, 0};
/* fm.cpp:30 four-letter classes */
static const int four_letter_map[20] = {1, 3, 2, 2, 4, 2, 2, 1, 3, 4, 4, 3, 4, 4, 1, 1, 1, 4, 4, 4};

enum { FM_CA = 1, FM_CB = 2 };

typedef struct or_mem_s {
  int pos, len;
  double weight;
  char *se;
  double *rf0, *rf1; /* len*len: rf0 CA-CA (i<j) / CB-CB (i>j), rf1 CA(i)-CB(j) */
  double *vmf;       /* len*len angles between the CA->CB vectors (vector fragment memory), or NULL */
} or_mem;

typedef struct {
  int nseq_cl, nres_cl, frag_resty;
  int i_sep[10];
  double *gamma;
} or_gamma;

struct or_ctx {
  /* flags (fb.cpp:124-130) */
  int abc_flag, chain_flag, shake_flag, chi_flag, rama_flag, rama_p_flag;
  int ssweight_flag, dssp_hdrgn_flag, p_ap_flag, water_flag, burial_flag;
  int helix_flag, frag_mem_flag;
  double epsilon;
  double k_chain[3], r_ncb0, r_cpcb0, r_ncp0;
  double k_shake, r_sh1, r_sh2, r_sh3;
  double k_chi, chi0, k_rama;
  int n_rama_par, n_rama_p_par;
  int ssweight[12];
  double w[12], sigma[12], phiw[12], phi0[12], psiw[12], psi0[12];
  double an, bn, cn, ap, bp, cp, ah, bh, ch;
  /* DSSP */
  double k_dssp, hbscl[4][9], sigma_HO, sigma_NO, HO_zero, NO_zero;
  double dssp_hdrgn_cut, dssp_hdrgn_cut_sq, pref[2], d_nu0;
  double dssp_nu_cut1_sq, dssp_nu_cut2_sq, sigma_HO_sqinv, sigma_NO_sqinv;
  double m_anti_HB[20][20][2], m_anti_NHB[20][20][2], m_para_HB[20][20][2];
  double m_para_one[20], m_anti_one[20];
  /* P_AP */
  double k_global_P_AP, k_betapred_P_AP, k_P_AP[3], P_AP_cut, P_AP_pref, pap_cutoff_sq;
  int i_med_min, i_med_max, i_diff_P_AP;
  double pap_drmax_sq, pap_drmin_sq;
  /* water / burial */
  double k_water, water_kappa, water_kappa_sigma, treshold;
  int contact_cutoff, n_wells, well_flag[5];
  double well_r_min[5], well_r_max[5];
  double w_rmin_theta[5], w_rmax_theta[5], w_rmin_theta_sq[5], w_rmax_theta_sq[5];
  double water_gamma[5][20][20][2];
  double k_burial, burial_kappa, burial_ro_min[3], burial_ro_max[3];
  double burial_gamma[20][3];
  /* helix */
  double k_helix, helix_gamma_p, helix_gamma_w, h4prob[20];
  double helix_kappa, helix_kappa_sigma, helix_treshold, helix_cutoff, helix_cutoff_sq;
  double helix_well_r_min[5], helix_well_r_max[5];
  double helix_sigma_HO, helix_sigma_NO, helix_HO_zero, helix_NO_zero;
  double helix_sigma_HO_sqinv, helix_sigma_NO_sqinv;
  int n_helix_wells, helix_i_diff, pro_accepter_flag;
  double h4prob_pro_accepter;
  double h_rmin_theta, h_rmax_theta, h_rmin_theta_sq, h_rmax_theta_sq;
  /* fragment memory */
  double k_frag_mem, fm_sigma_exp;
  char frag_mems_file[256], fm_gamma_file[256];
  or_gamma fmg;
  int n_mems;
  or_mem *mems;
  int *ilen_fm_map, **frag_mem_map;
  /* fragment-memory table mode ([Fragment_Memory_Table], fb.cpp:310-320):
     per (i, j - i - minSep, atom pair k) a curve of energy and force on
     the grid r = tb_rmin + ir tb_dr, ir < tb_size; NULL where no memory
     contributes (fb.cpp:997-1017, 4238-4323) */
  int frag_mem_tb_flag, tb_size, tb_nbrs, fm_use_table_flag;
  double tb_rmin, tb_rmax, tb_dr, frag_table_well_width;
  double **tb_e, **tb_f;
  char last_err[256];
  double pair_list_cutoff;
  double list_skin;                     /* neighbour-list skin of the list-limited terms (or_set_list_skin) */
  /* coeff-toggled terms: AMH-Go (fb.cpp:297-303), contact restraints
     (321-327), vector fragment memory (328-333), solvent-separated barrier
     (334-345), membrane (346-358), phosphorylation (412-421), Debye-Hueckel
     (466-478) */
  int amh_go_flag, cont_rest_flag, vec_frag_mem_flag, ssb_flag, memb_flag, phosph_flag, huckel_flag;
  double k_amh_go, amh_go_p, amh_go_rc, amh_go_norm;
  int frustration_censoring_flag;
  or_gamma amhg;
  or_mem amh_mem;                       /* amh-go.gro as one memory (pos 0, len n) */
  unsigned char *censor;                /* n*n, censoring flag 1 */
  double *rn_caca, *rn_cbcb, *rn_cacb;  /* n*n, censoring flag 2 */
  double *go_f;                         /* 2n*3 scratch: one atom's pair forces */
  double k_cont_rest, cr_sigma, cr_sigma_sq_inv, cr_dr_cutoff, cr_glob_cutoff_sq;
  char cr_file[256];
  int n_cr;
  int *cr_i1, *cr_i2, *cr_start;        /* entries sorted by (i1, i2, file order); start per i1 (n+1) */
  double *cr_w, *cr_r0;
  double k_vec_frag_mem, vfm_sigma, vfm_sigma_sq;
  double k_solventb1, ssb_rmin1, ssb_rmax1, k_solventb2, ssb_rmin2, ssb_rmax2, ssb_kappa, ssb_rshift[20];
  int ssb_ij_sep, ssb_rad_cor;
  double k_overall_memb, k_bin, memb_xo[3], memb_len, rho0_max, rho0_distor, g_memb[3][4];
  int memb_pore_type;
  int *z_res;
  double k_hypercharge;
  int n_phosph_res, phosph_res[20];
  int *phosph_map;
  double phosph_water_gamma[5][20][20][2];
  double k_PlusPlus, k_MinusMinus, k_PlusMinus, k_screening, screening_length;
  int debye_huckel_min_sep;
  double *charge;
  /* [Tertiary_Frustratometer] (fb.cpp:385-397): contact cutoff, decoys per
     contact, output interval, mode 0 configurational / 1 mutational / 2
     singleresidue; the configurational decoy statistics, computed once
     (fb.cpp:5099, 5342); the rand() stream the decoys draw from */
  int tert_frust_flag, tf_ndecoys, tf_freq, tf_mode, tf_done;
  double tf_cutoff, tf_stats[2];
  /* [Nmer_Frustratometer] (fb.cpp:398-412): n-mer size, contact cutoff,
     contacts needed, decoys, output interval, the tcl thresholds and
     neutral flag, the trap flags, trap width in sigmas, secondary-structure
     fraction, mode 0 pairwise / 1 singlenmer */
  int nmer_frust_flag, nf_size, nf_contacts, nf_ndecoys, nf_freq, nf_neutral, nf_trap, nf_draw_trap, nf_mode;
  double nf_cutoff, nf_min_thr, nf_high_thr, nf_trap_sigma, nf_ss_frac;
  /* [Fragment_Frustratometer] (fb.cpp:359-384): mode 0 shuffle / 1 read,
     the decoy library (shuffled memories or decoy structures), decoy
     calculations (shuffle), output interval, well width, sequence-
     separation scaling (read mode only; shuffle mode leaves them unset in
     the reference: width 1, no scaling here); the decoy map, and read
     mode's per-residue decoy statistics (computed once) */
  int frag_frust_flag, ff_mode, ff_ncalc, ff_freq, ff_seqsep, ff_norm, ff_gen_done;
  double ff_ww, ff_seqsep_gamma;
  char ff_file[256];
  struct or_mem_s *ff_mems;
  int ff_nm;
  int *ff_ilen, **ff_map;
  double *ff_mean, *ff_var;
  uint32_t rng[34];
  int rng_k;
  /* system */
  int n;
  char *se;
  int nch, ch_len[MAXCH], ch_pos[MAXCH];
  int *chain_no; /* 1-based chain tag per residue */
  double *aps[12];
  /* excluded volume pair style (pev.cpp) */
  int ev_p;
  double ev_lambda[6][6], ev_cut_short[6][6], ev_cut_long[6][6];
  /* bonds (LAMMPS harmonic, deck of CoordinatesToWorkLammpsDataFile.py:289-294) */
  double bond_k, bond_r0[5];
  /* scratch */
  double *xn, *xh, *xcp;
  double *water_ro, *helix_ro, *w_sig, *w_sig_prd, *h_sig, *h_sig_prd;
  double *water_xi, *helix_xi_1, *helix_xi_2, *burial_force;
  /* current call */
  const double *xca, *xcb, *xo;
  double *fca, *fcb, *fo;
  double *energy;
  unsigned mask;
  /* branch counters (or_branch_counts): which selection branches ran */
  long long branch[OR_N_BRANCH];
};

/* ------------------------------------------------------------------ */
/* small helpers                                                        */

static void set_err(char *err, int errlen, const char *fmt, ...)
{
  if (!err || errlen <= 0) return;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(err, (size_t)errlen, fmt, ap);
  va_end(ap);
}

static char *read_file(const char *path, long *len)
{
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *buf = (char *)malloc((size_t)n + 1);
  if (!buf) { fclose(f); return NULL; }
  size_t got = fread(buf, 1, (size_t)n, f);
  buf[got] = '\0';
  fclose(f);
  if (len) *len = (long)got;
  return buf;
}

static void join_path(char *out, size_t n, const char *dir, const char *name)
{
  if (name[0] == '/' || !dir || !dir[0]) snprintf(out, n, "%s", name);
  else snprintf(out, n, "%s/%s", dir, name);
}

/* istream-like extraction over a text buffer: skip whitespace, then take
 * the longest prefix valid for the requested type (operator>> semantics). */
typedef struct { const char *p; int fail; int eof; } tokin;

static void tk_skip(tokin *t)
{
  while (*t->p && isspace((unsigned char)*t->p)) t->p++;
  if (!*t->p) t->eof = 1;
}

static int tk_str(tokin *t, char *out, size_t n)
{
  if (t->fail) return 0;
  tk_skip(t);
  if (!*t->p) { t->fail = 1; return 0; }
  size_t k = 0;
  while (*t->p && !isspace((unsigned char)*t->p)) {
    if (k + 1 < n) out[k++] = *t->p;
    t->p++;
  }
  out[k] = '\0';
  if (!*t->p) t->eof = 1;
  return 1;
}

static double tk_double(tokin *t)
{
  if (t->fail) return 0.0;
  tk_skip(t);
  if (!*t->p) { t->fail = 1; return 0.0; }
  char *end;
  double v = strtod(t->p, &end);
  if (end == t->p) { t->fail = 1; return 0.0; }
  t->p = end;
  if (!*t->p) t->eof = 1;
  return v;
}

static int tk_int(tokin *t)
{
  if (t->fail) return 0;
  tk_skip(t);
  if (!*t->p) { t->fail = 1; return 0; }
  char *end;
  long v = strtol(t->p, &end, 10);
  if (end == t->p) { t->fail = 1; return 0; }
  t->p = end;
  if (!*t->p) t->eof = 1;
  return (int)v;
}

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void cross3(const double *a, const double *b, double *c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

#define XCA(i) (c->xca + 3 * (i))
#define XCB(i) (c->xcb + 3 * (i))
#define XO(i) (c->xo + 3 * (i))
#define XN(i) (c->xn + 3 * (i))
#define XH(i) (c->xh + 3 * (i))
#define XCP(i) (c->xcp + 3 * (i))
#define FCA(i) (c->fca + 3 * (i))
#define FCB(i) (c->fcb + 3 * (i))
#define FO(i) (c->fo + 3 * (i))
#define ON(t) (c->mask & (1u << (t)))

static inline void fadd(double *f, double s, const double *v)
{
  f[0] += s * v[0];
  f[1] += s * v[1];
  f[2] += s * v[2];
}

/* letters outside A..Z index out of se_map in the reference (undefined);
 * they map to type 0 here */
static inline int restype(char ch) { return (ch >= 'A' && ch <= 'Z') ? se_map[ch - 'A'] : 0; }

/* isFirst / isLast, fb.cpp:1570-1583 (1-based chain positions, 0-based index) */
static inline int is_first(const or_ctx *c, int i) { return c->ch_pos[c->chain_no[i] - 1] == i + 1; }
static inline int is_last(const or_ctx *c, int i)
{
  int ch = c->chain_no[i] - 1;
  return c->ch_pos[ch] + c->ch_len[ch] - 1 == i + 1;
}

/* ------------------------------------------------------------------ */
/* Gamma_Array (fm.cpp:275-519)                                         */

static int gm_index_array(or_gamma *g, const char *resty, int *a)
{
  if (strcmp(resty, "ALL") == 0) {
    for (int i = 0; i < g->nres_cl; ++i) a[i] = i;
    return g->nres_cl;
  } else if (strlen(resty) == 1) {
    int code = resty[0] - 'A';
    if (code < 0 || code >= 26) return -1;
    if (se_map[code] == 0 && resty[0] != 'A') return -1;
    if (g->nres_cl < 20) return -2;
    a[0] = se_map[code];
    return 1;
  } else if (strlen(resty) == 3) {
    int cl;
    if (g->nres_cl < 4) return -2;
    if (strcmp(resty, "SHL") == 0) cl = 1;
    else if (strcmp(resty, "AHL") == 0) cl = 2;
    else if (strcmp(resty, "HPB") == 0) cl = 4;
    else if (strcmp(resty, "BAS") == 0) cl = 3;
    else return -1;
    if (g->nres_cl == 4) { a[0] = cl - 1; return 1; }
    int n = 0;
    for (int i = 0; i < 20; ++i) if (four_letter_map[i] == cl) a[n++] = i;
    return n;
  }
  return -1;
}

static int gm_load(or_gamma *g, const char *path, char *err, int errlen)
{
  memset(g, 0, sizeof(*g));
  FILE *f = fopen(path, "r");
  if (!f) { set_err(err, errlen, "Fragment_Memory: Cannot read gamma file"); return -1; }
  char line[100];
  int iline = 0, frag_set = 0;
  long datapos = 0;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int empty = 1;
    for (char *q = line; *q; ++q) if (*q != ' ' && *q != '\t' && *q != '\n') { empty = 0; break; }
    if (empty) continue;
    if (iline == 0) {
      int ns = 0;
      char *st = strtok(line, " \t");
      while (st) {
        ns++;
        if (strcmp(st, "inf") == 0) { g->i_sep[ns - 1] = -1; break; }
        g->i_sep[ns - 1] = atoi(st);
        st = strtok(NULL, " \t\n");
      }
      g->nseq_cl = ns - 1;
      if (g->nseq_cl == 0) { fclose(f); set_err(err, errlen, "Fragment_Memory: Wrong definition of sequance separation classes"); return -1; }
      datapos = ftell(f);
    } else {
      char buf[8][64];
      int nbuf = 0;
      char *st = strtok(line, " \t\n");
      while (st && nbuf < 8) { snprintf(buf[nbuf++], 64, "%s", st); st = strtok(NULL, " \t\n"); }
      if (nbuf != 4 && nbuf != 6) { fclose(f); set_err(err, errlen, "Fragment_Memory: Incorrect entery in gamma file"); return -1; }
      if (!frag_set) { g->frag_resty = (nbuf == 6); frag_set = 1; }
      else if ((nbuf == 4 && g->frag_resty) || (nbuf == 6 && !g->frag_resty)) {
        fclose(f); set_err(err, errlen, "Fragment_Memory: Incorrect entery in gamma file"); return -1;
      }
      for (int i = 0; i < nbuf - 2; i++) {
        size_t bl = strlen(buf[i]);
        if (bl != 1 && bl != 3) { fclose(f); set_err(err, errlen, "Fragment_Memory: Incorrect entery in gamma file"); return -1; }
        if (strcmp(buf[i], "ALL") == 0) { if (g->nres_cl < 1) g->nres_cl = 1; }
        else if (bl == 3) { if (g->nres_cl < 4) g->nres_cl = 4; }
        else { if (g->nres_cl < 20) g->nres_cl = 20; }
      }
    }
    iline++;
  }
  if (g->nres_cl == 0) { fclose(f); set_err(err, errlen, "Fragment_Memory: Incorrect entery in gamma file"); return -1; }
  int r = g->nres_cl;
  long ng = g->frag_resty ? (long)g->nseq_cl * r * r * r * r : (long)g->nseq_cl * r * r;
  g->gamma = (double *)calloc((size_t)ng, sizeof(double));
  fseek(f, datapos, SEEK_SET);
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int empty = 1;
    for (char *q = line; *q; ++q) if (*q != ' ' && *q != '\t' && *q != '\n') { empty = 0; break; }
    if (empty) continue;
    char *t1 = strtok(line, " \t\n");
    char *t2 = strtok(NULL, " \t\n");
    if (!g->frag_resty) {
      int cl = atoi(strtok(NULL, " \t\n"));
      double gm = atof(strtok(NULL, " \t\n"));
      if (cl > g->nseq_cl) { fclose(f); set_err(err, errlen, "Fragment_Memory: Wrong sequance separation class tag"); return -1; }
      int is[20], js[20];
      int in = gm_index_array(g, t1, is), jn = gm_index_array(g, t2, js);
      if (in < 0 || jn < 0) {
        fclose(f);
        set_err(err, errlen, (in == -2 || jn == -2) ? "Fragment_Memory: Cannot build gamma array" : "Fragment_Memory: Incorrect entery in gamma file");
        return -1;
      }
      for (int i = 0; i < in; ++i)
        for (int j = 0; j < jn; ++j) {
          g->gamma[(cl - 1) * r * r + is[i] * r + js[j]] = gm;
          if (is[i] != js[j]) g->gamma[(cl - 1) * r * r + js[j] * r + is[i]] = gm;
        }
    } else {
      char *t3 = strtok(NULL, " \t\n");
      (void)strtok(NULL, " \t\n"); /* jfresty: the reference passes ifresty twice (fm.cpp:371) */
      int cl = atoi(strtok(NULL, " \t\n"));
      double gm = atof(strtok(NULL, " \t\n"));
      if (cl > g->nseq_cl) { fclose(f); set_err(err, errlen, "Fragment_Memory: Cannot build gamma array"); return -1; }
      int is[20], js[20], ks[20], ls[20];
      int in = gm_index_array(g, t1, is), jn = gm_index_array(g, t2, js);
      int kn = gm_index_array(g, t3, ks), ln = gm_index_array(g, t3, ls);
      if (in < 0 || jn < 0 || kn < 0 || ln < 0) { fclose(f); set_err(err, errlen, "Fragment_Memory: Incorrect entery in gamma file"); return -1; }
      long b = (long)(cl - 1) * r * r * r * r;
      for (int i = 0; i < in; ++i)
        for (int j = 0; j < jn; ++j)
          for (int k = 0; k < kn; ++k)
            for (int l = 0; l < ln; ++l) {
              g->gamma[b + is[i] * r * r * r + js[j] * r * r + ks[k] * r + ls[l]] = gm;
              if (is[i] != js[j]) g->gamma[b + js[j] * r * r * r + is[i] * r * r + ks[k] * r + ls[l]] = gm;
              if (ks[k] != ls[l]) g->gamma[b + is[i] * r * r * r + js[j] * r * r + ls[l] * r + ks[k]] = gm;
              if (is[i] != js[j] && ks[k] != ls[l]) g->gamma[b + js[j] * r * r * r + is[i] * r * r + ls[l] * r + ks[k]] = gm;
            }
    }
  }
  fclose(f);
  return 0;
}

/* Gamma_Array::getGamma(ires_type, jres_type, ires, jres), fm.cpp:400-424 */
static double gm_get2(const or_gamma *g, int it, int jt, int ires, int jres)
{
  int dij = abs(ires - jres);
  if (dij < g->i_sep[0] || (g->i_sep[g->nseq_cl] != -1 && dij > g->i_sep[g->nseq_cl])) return 0.0;
  int seq_cl;
  for (seq_cl = 1; seq_cl < g->nseq_cl && dij >= g->i_sep[seq_cl]; seq_cl++) {}
  int r = g->nres_cl;
  if (r == 1) return g->gamma[seq_cl - 1];
  int icl = it, jcl = jt;
  if (r == 4) { icl = four_letter_map[it] - 1; jcl = four_letter_map[jt] - 1; }
  return g->gamma[(seq_cl - 1) * r * r + icl * r + jcl];
}

/* getGamma with fragment residue types, fm.cpp:426-454 */
static double gm_get4(const or_gamma *g, int it, int jt, int ift, int jft, int ires, int jres)
{
  int dij = abs(ires - jres);
  if (dij < g->i_sep[0] || (g->i_sep[g->nseq_cl] != -1 && dij > g->i_sep[g->nseq_cl])) return 0.0;
  int seq_cl;
  for (seq_cl = 1; seq_cl < g->nseq_cl && dij >= g->i_sep[seq_cl]; seq_cl++) {}
  int r = g->nres_cl;
  if (r == 1) return g->gamma[seq_cl - 1];
  int a = it, b = jt, cc = ift, d = jft;
  if (r == 4) { a = four_letter_map[it] - 1; b = four_letter_map[jt] - 1; cc = four_letter_map[ift] - 1; d = four_letter_map[jft] - 1; }
  return g->gamma[(long)(seq_cl - 1) * r * r * r * r + a * r * r * r + b * r * r + cc * r + d];
}

/* ------------------------------------------------------------------ */
/* Fragment_Memory (.gro reader, fm.cpp:34-176) and read_mems (fb.cpp:1966-2042) */

static char three_to_one(const char *s)
{
  static const char *names[20] = {"ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
                                  "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL"};
  static const char one[20] = {'A', 'R', 'N', 'D', 'C', 'Q', 'E', 'G', 'H', 'I', 'L', 'K', 'M', 'F', 'P', 'S', 'T', 'W', 'Y', 'V'};
  if (strlen(s) == 3)
    for (int i = 0; i < 20; ++i) if (strcmp(s, names[i]) == 0) return one[i];
  return '?';
}

/* The atom records of one .gro file in file order (the fields fm.cpp:60-110
 * reads), parsed once per file and shared by every memory taken from it. */
typedef struct {
  int ires, kind; /* kind: 1 CA, 2 CB, 0 any other atom */
  char one;       /* one-letter residue name ('?' when unknown, fm.cpp:270) */
  double x, y, z; /* nm */
} gro_rec;
typedef struct {
  char path[1024];
  int ok, n;
  gro_rec *r;
} gro_file;

static void gro_parse(const char *path, gro_file *g)
{
  memset(g, 0, sizeof(*g));
  snprintf(g->path, sizeof g->path, "%s", path);
  FILE *f = fopen(path, "r");
  if (!f) return;
  g->ok = 1;
  char buff[201];
  int nAtoms = 0;
  if (!fgets(buff, 200, f)) nAtoms = 0;
  if (fscanf(f, "%d", &nAtoms) != 1) nAtoms = 0;
  if (!fgets(buff, 200, f)) nAtoms = 0;
  g->r = (gro_rec *)calloc((size_t)(nAtoms > 0 ? nAtoms : 1), sizeof(gro_rec));
  for (int i = 0; i < nAtoms; ++i) {
    int ires, iatom;
    char resty[64], atomty[64];
    double x, y, z;
    if (fscanf(f, "%d %63s %63s %d %lf %lf %lf", &ires, resty, atomty, &iatom, &x, &y, &z) != 7) break;
    gro_rec *q = &g->r[g->n++];
    q->ires = ires;
    q->kind = strcmp(atomty, "CA") == 0 ? 1 : (strcmp(atomty, "CB") == 0 ? 2 : 0);
    q->one = three_to_one(resty);
    q->x = x; q->y = y; q->z = z;
  }
  fclose(f);
}

static int mem_load(or_mem *m, int pos, int fpos, int len, double weight, const gro_file *g, int vec)
{
  memset(m, 0, sizeof(*m));
  m->pos = pos; m->len = len; m->weight = weight;
  m->se = (char *)calloc((size_t)len + 1, 1);
  if (!g->ok) return 1; /* ERR_FILE */
  double *xca = (double *)calloc((size_t)len * 3, sizeof(double));
  double *xcb = (double *)calloc((size_t)len * 3, sizeof(double));
  int nca = 0, ncb = 0;
  for (int i = 0; i < g->n; ++i) {
    const gro_rec *q = &g->r[i];
    int ires = q->ires;
    if (ires > fpos + len) break;
    if (ires > fpos && ires <= fpos + len) {
      ires -= fpos + 1;
      double x = q->x * 10, y = q->y * 10, z = q->z * 10;
      if (q->kind == 1) {
        if (ires >= len || nca >= len) { free(xca); free(xcb); return 2; }
        m->se[ires] = q->one;
        if (m->se[ires] == '-') { free(xca); free(xcb); return 3; }
        xca[3 * ires] = x; xca[3 * ires + 1] = y; xca[3 * ires + 2] = z;
        nca++;
      }
      if (q->kind == 2) {
        if (ires >= len || ncb >= len) { free(xca); free(xcb); return 2; }
        xcb[3 * ires] = x; xcb[3 * ires + 1] = y; xcb[3 * ires + 2] = z;
        ncb++;
      }
    }
  }
  if (nca != len) { free(xca); free(xcb); return 2; }
  m->rf0 = (double *)calloc((size_t)len * len, sizeof(double));
  m->rf1 = (double *)calloc((size_t)len * len, sizeof(double));
  for (int i = 0; i < len; ++i)
    for (int j = 0; j < len; ++j) {
      const double *a, *b;
      if (i < j) { a = xca + 3 * i; b = xca + 3 * j; }
      else if (i > j) { a = xcb + 3 * i; b = xcb + 3 * j; }
      else { a = b = NULL; }
      if (a) {
        double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
        m->rf0[i * len + j] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
      }
      const double *p = xca + 3 * i, *q = xcb + 3 * j;
      double e0 = p[0] - q[0], e1 = p[1] - q[1], e2 = p[2] - q[2];
      m->rf1[i * len + j] = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    }
  if (vec) { /* Fragment_Memory::VM, fm.cpp:166, 239-243 */
    m->vmf = (double *)calloc((size_t)len * len, sizeof(double));
    for (int i = 0; i < len; ++i)
      for (int j = 0; j < len; ++j) {
        const double *a1 = xca + 3 * i, *b1 = xcb + 3 * i, *a2 = xca + 3 * j, *b2 = xcb + 3 * j;
        double u[3] = {b1[0] - a1[0], b1[1] - a1[1], b1[2] - a1[2]}, v[3] = {b2[0] - a2[0], b2[1] - a2[1], b2[2] - a2[2]};
        m->vmf[i * len + j] = acos(dot3(u, v) / (sqrt(dot3(u, u)) * sqrt(dot3(v, v))));
      }
  }
  free(xca); free(xcb);
  return 0;
}

/* Fragment_Memory::Rf, fm.cpp:190-202 */
static double mem_rf(const or_mem *m, int ires, int iatom, int jres, int jatom)
{
  ires -= m->pos; jres -= m->pos;
  if (ires < 0 || ires >= m->len || jres < 0 || jres >= m->len) return NAN;
  int lo = ires < jres ? ires : jres, hi = ires < jres ? jres : ires;
  if (iatom == FM_CA && jatom == FM_CA) return m->rf0[lo * m->len + hi];
  if (iatom == FM_CB && jatom == FM_CB) return m->rf0[hi * m->len + lo];
  return iatom == FM_CA ? m->rf1[ires * m->len + jres] : m->rf1[jres * m->len + ires];
}

static char *trim(char *s)
{
  while (isspace((unsigned char)*s)) s++;
  size_t n = strlen(s);
  if (n == 0) return s;
  char *b = s + n;
  while (b > s && isspace((unsigned char)*--b)) {}
  if (!isspace((unsigned char)*b)) b[1] = '\0';
  else *b = '\0';
  return s;
}

static int read_mems(or_ctx *c, const char *workdir, const char *mems_file, or_mem **mems_out, int *nm_out, int vec,
                     char *err, int errlen)
{
  char path[1024];
  join_path(path, sizeof path, workdir, mems_file);
  FILE *f = fopen(path, "r");
  if (!f) { set_err(err, errlen, "Fragment_Memory: Error opening mem file"); return -1; }
  char ln[500];
  int state = 0; /* 0 none, 1 target, 2 mems */
  *nm_out = 0;
  *mems_out = NULL;
  int cap = 0;
  int ngro = 0, capgro = 0;
  gro_file *gros = NULL;
  int rc = 0;
  while (fgets(ln, sizeof ln, f)) {
    char *line = trim(ln);
    if (line[0] == '#') continue;
    if (line[0] == '[') state = 0;
    int empty = 1;
    for (char *q = line; *q; ++q) if (*q != ' ' && *q != '\t' && *q != '\n') { empty = 0; break; }
    if (empty) { state = 0; continue; }
    if (state == 2) {
      char *str[6];
      int nstr = 0;
      str[nstr] = strtok(line, " \t\n");
      while (str[nstr] != NULL) {
        nstr++;
        if (nstr > 5) break;
        str[nstr] = strtok(NULL, " \t\n");
      }
      if (nstr != 5) { set_err(err, errlen, "Fragment_Memory: Error reading mem file"); rc = -1; break; }
      int tpos = atoi(str[1]) - 1, fpos = atoi(str[2]) - 1, len = atoi(str[3]);
      double weight = atof(str[4]);
      if (*nm_out == cap) { cap = cap ? 2 * cap : 64; *mems_out = (or_mem *)realloc(*mems_out, (size_t)cap * sizeof(or_mem)); }
      char gro[1024];
      join_path(gro, sizeof gro, workdir, str[0]);
      int gi = 0;
      while (gi < ngro && strcmp(gros[gi].path, gro) != 0) ++gi;
      if (gi == ngro) {
        if (ngro == capgro) { capgro = capgro ? 2 * capgro : 16; gros = (gro_file *)realloc(gros, (size_t)capgro * sizeof(gro_file)); }
        gro_parse(gro, &gros[ngro++]);
      }
      int e = mem_load(&(*mems_out)[*nm_out], tpos, fpos, len, weight, &gros[gi], vec);
      (*nm_out)++;
      if (e) { set_err(err, errlen, "read_mems: Fragment_Memory: Error reading memory fragment"); rc = -1; break; }
      if (tpos + len > c->n) { set_err(err, errlen, "read_mems: Fragment_Memory: Incorrectly defined memory fragment"); rc = -1; break; }
    } else if (state == 0) {
      if (strcmp(line, "[Target]") == 0) state = 1;
      else if (strcmp(line, "[Memories]") == 0) state = 2;
    }
  }
  fclose(f);
  for (int k = 0; k < ngro; ++k) free(gros[k].r);
  free(gros);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Constructor: coeff parse (fb.cpp:165-497), scaling (502-516), side   */
/* files (538-791), pair_list_cutoff (1048-1071)                        */

/* the decoy memory map at the decoys' current positions (fb.cpp:812-834,
 * 4588-4610): residue i lists the decoys with pos <= i < pos + len - minSep */
static int ff_build_map(or_ctx *c)
{
  int n = c->n, min_sep = c->fmg.i_sep[0];
  for (int i = 0; i < n; ++i) { free(c->ff_map[i]); c->ff_map[i] = NULL; c->ff_ilen[i] = 0; }
  for (int k = 0; k < c->ff_nm; ++k) {
    int pos = c->ff_mems[k].pos, len = c->ff_mems[k].len;
    if (pos + len > n) return -1;
    for (int i = pos; i < pos + len - min_sep; ++i) {
      c->ff_ilen[i]++;
      c->ff_map[i] = (int *)realloc(c->ff_map[i], (size_t)c->ff_ilen[i] * sizeof(int));
      c->ff_map[i][c->ff_ilen[i] - 1] = k;
    }
  }
  return 0;
}

static double calc_exp_helix_cutoff(const or_ctx *c) /* fb.cpp:1074-1104 */
{
  double r0[3] = {3.77, 2.5, 2.41}, r1[3], r2[3];
  double an = c->an, bn = c->bn, cn = c->cn, ah = c->ah, bh = c->bh, ch = c->ch;
  r1[0] = sqrt(pow(an * r0[0], 2) + pow(cn * r0[1], 2) + an * cn * (r0[0] * r0[0] + r0[1] * r0[1] - r0[2] * r0[2]));
  r1[1] = sqrt(pow(an * r0[2], 2) + pow(bn * r0[1], 2) + an * bn * (r0[2] * r0[2] + r0[1] * r0[1] - r0[0] * r0[0]));
  r1[2] = sqrt(pow(bn * r0[0], 2) + pow(cn * r0[2], 2) + bn * cn * (r0[0] * r0[0] + r0[2] * r0[2] - r0[1] * r0[1]));
  r2[0] = sqrt(pow(ah * r0[0], 2) + pow(ch * r0[1], 2) + ah * ch * (r0[0] * r0[0] + r0[1] * r0[1] - r0[2] * r0[2]));
  r2[1] = sqrt(pow(ah * r0[2], 2) + pow(bh * r0[1], 2) + ah * bh * (r0[2] * r0[2] + r0[1] * r0[1] - r0[0] * r0[0]));
  r2[2] = sqrt(pow(bh * r0[0], 2) + pow(ch * r0[2], 2) + bh * ch * (r0[0] * r0[0] + r0[2] * r0[2] - r0[1] * r0[1]));
  double r1m = fmax(r1[0], fmax(r1[1], r1[2])), r2m = fmax(r2[0], fmax(r2[1], r2[2]));
  r1m = 1.1 * r1m + c->helix_NO_zero + c->helix_sigma_NO * sqrt(6.0 * log(10.0));
  r2m = 1.1 * r2m + c->helix_HO_zero + c->helix_sigma_HO * sqrt(6.0 * log(10.0));
  return fmax(r1m, r2m);
}

static int parse_coeff(or_ctx *c, const char *path, char *err, int errlen)
{
  char *text = read_file(path, NULL);
  if (!text) { set_err(err, errlen, "Coefficient file was not found!"); return -1; }
  tokin in = {text, 0, 0};
  char sec[256], sbuf[256];
  int j;
  while (!in.eof && !in.fail) {
    sec[0] = '\0';
    tk_str(&in, sec, sizeof sec);
    if (strcmp(sec, "[ABC]") == 0) {
      c->abc_flag = 1;
      c->an = tk_double(&in); c->bn = tk_double(&in); c->cn = tk_double(&in);
      c->ap = tk_double(&in); c->bp = tk_double(&in); c->cp = tk_double(&in);
      c->ah = tk_double(&in); c->bh = tk_double(&in); c->ch = tk_double(&in);
    } else if (strcmp(sec, "[Chain]") == 0) {
      c->chain_flag = 1;
      for (j = 0; j < 3; j++) c->k_chain[j] = tk_double(&in);
      c->r_ncb0 = tk_double(&in); c->r_cpcb0 = tk_double(&in); c->r_ncp0 = tk_double(&in);
    } else if (strcmp(sec, "[Shake]") == 0) {
      c->shake_flag = 1;
      c->k_shake = tk_double(&in); c->r_sh1 = tk_double(&in); c->r_sh2 = tk_double(&in); c->r_sh3 = tk_double(&in);
    } else if (strcmp(sec, "[Chi]") == 0) {
      c->chi_flag = 1;
      c->k_chi = tk_double(&in); c->chi0 = tk_double(&in);
    } else if (strcmp(sec, "[Excluded]") == 0 || strcmp(sec, "[Excluded_R6]") == 0) {
      /* production build: parsed, used only under DEBUGFORCES */
      for (j = 0; j < 4; j++) tk_double(&in);
    } else if (strcmp(sec, "[Excluded_P]") == 0) {
      tk_int(&in);
      for (j = 0; j < 4; j++) tk_double(&in);
    } else if (strcmp(sec, "[Rama]") == 0) {
      c->rama_flag = 1;
      c->k_rama = tk_double(&in);
      c->n_rama_par = tk_int(&in);
      if (c->n_rama_par > 6 || c->n_rama_par < 0) { free(text); set_err(err, errlen, "Rama: too many wells"); return -1; }
      for (j = 0; j < c->n_rama_par; j++) {
        c->w[j] = tk_double(&in); c->sigma[j] = tk_double(&in); c->phiw[j] = tk_double(&in);
        c->phi0[j] = tk_double(&in); c->psiw[j] = tk_double(&in); c->psi0[j] = tk_double(&in);
        c->phiw[j] *= c->sigma[j];
        c->psiw[j] *= c->sigma[j];
      }
    } else if (strcmp(sec, "[Rama_P]") == 0) {
      c->rama_p_flag = 1;
      c->n_rama_p_par = tk_int(&in);
      if (c->n_rama_p_par > 6 || c->n_rama_p_par < 0) { free(text); set_err(err, errlen, "Rama_P: too many wells"); return -1; }
      for (j = 0; j < c->n_rama_p_par; j++) {
        int k = j + 6;
        c->w[k] = tk_double(&in); c->sigma[k] = tk_double(&in); c->phiw[k] = tk_double(&in);
        c->phi0[k] = tk_double(&in); c->psiw[k] = tk_double(&in); c->psi0[k] = tk_double(&in);
        c->phiw[k] *= c->sigma[k];
        c->psiw[k] *= c->sigma[k];
      }
    } else if (strcmp(sec, "[SSWeight]") == 0) {
      c->ssweight_flag = 1;
      for (j = 0; j < 12; ++j) c->ssweight[j] = tk_int(&in) != 0;
    } else if (strcmp(sec, "[Dssp_Hdrgn]") == 0) {
      c->dssp_hdrgn_flag = 1;
      c->k_dssp = tk_double(&in);
      c->hbscl[0][0] = tk_double(&in); c->hbscl[0][1] = tk_double(&in);
      for (j = 0; j < 7; ++j) c->hbscl[1][j] = tk_double(&in);
      for (j = 0; j < 9; ++j) c->hbscl[2][j] = tk_double(&in);
      for (j = 0; j < 9; ++j) c->hbscl[3][j] = tk_double(&in);
      c->sigma_HO = tk_double(&in); c->sigma_NO = tk_double(&in);
      c->HO_zero = tk_double(&in); c->NO_zero = tk_double(&in);
      c->dssp_hdrgn_cut = tk_double(&in);
      c->pref[0] = tk_double(&in); c->pref[1] = tk_double(&in);
      c->d_nu0 = tk_double(&in);
      c->dssp_hdrgn_cut_sq = c->dssp_hdrgn_cut * c->dssp_hdrgn_cut;
      c->dssp_nu_cut1_sq = pow(c->d_nu0 + atanh(2.0 * DSSP_NU_DELTA - 1.0) / c->pref[0], 2);
      c->dssp_nu_cut2_sq = pow(c->d_nu0 + atanh(2.0 * DSSP_NU_DELTA - 1.0) / c->pref[1], 2);
      c->sigma_HO_sqinv = 1.0 / (c->sigma_HO * c->sigma_HO);
      c->sigma_NO_sqinv = 1.0 / (c->sigma_NO * c->sigma_NO);
    } else if (strcmp(sec, "[P_AP]") == 0) {
      c->p_ap_flag = 1;
      c->k_global_P_AP = tk_double(&in);
      c->k_betapred_P_AP = tk_double(&in);
      for (j = 0; j < 3; j++) c->k_P_AP[j] = tk_double(&in);
      c->P_AP_cut = tk_double(&in);
      c->P_AP_pref = tk_double(&in);
      c->i_med_min = tk_int(&in); c->i_med_max = tk_int(&in);
      c->i_diff_P_AP = tk_int(&in);
      c->pap_cutoff_sq = pow(c->P_AP_cut + atanh(1.0 - 2.0 * PAP_DELTA) / c->P_AP_pref, 2);
    } else if (strcmp(sec, "[Water]") == 0) {
      c->water_flag = 1;
      c->k_water = tk_double(&in);
      c->water_kappa = tk_double(&in); c->water_kappa_sigma = tk_double(&in);
      c->treshold = tk_double(&in);
      c->contact_cutoff = tk_int(&in);
      c->n_wells = tk_int(&in);
      if (c->n_wells > 5 || c->n_wells < 0) { free(text); set_err(err, errlen, "Water: too many wells"); return -1; }
      for (j = 0; j < c->n_wells; ++j) {
        c->well_r_min[j] = tk_double(&in); c->well_r_max[j] = tk_double(&in); c->well_flag[j] = tk_int(&in);
      }
    } else if (strcmp(sec, "[Burial]") == 0) {
      c->burial_flag = 1;
      c->k_burial = tk_double(&in);
      c->burial_kappa = tk_double(&in);
      for (j = 0; j < 3; j++) { c->burial_ro_min[j] = tk_double(&in); c->burial_ro_max[j] = tk_double(&in); }
    } else if (strcmp(sec, "[Helix]") == 0) {
      c->helix_flag = 1;
      c->k_helix = tk_double(&in);
      c->helix_gamma_p = tk_double(&in); c->helix_gamma_w = tk_double(&in);
      c->helix_kappa = tk_double(&in); c->helix_kappa_sigma = tk_double(&in);
      c->helix_treshold = tk_double(&in);
      c->helix_i_diff = tk_int(&in);
      c->helix_cutoff = tk_double(&in);
      c->helix_cutoff_sq = c->helix_cutoff * c->helix_cutoff;
      c->n_helix_wells = 1;
      c->helix_well_r_min[0] = tk_double(&in); c->helix_well_r_max[0] = tk_double(&in);
      for (j = 0; j < 20; ++j) c->h4prob[j] = tk_double(&in);
      c->pro_accepter_flag = tk_int(&in); c->h4prob_pro_accepter = tk_double(&in);
      c->helix_sigma_HO = tk_double(&in); c->helix_sigma_NO = tk_double(&in);
      c->helix_HO_zero = tk_double(&in); c->helix_NO_zero = tk_double(&in);
      c->helix_sigma_HO_sqinv = 1.0 / (c->helix_sigma_HO * c->helix_sigma_HO);
      c->helix_sigma_NO_sqinv = 1.0 / (c->helix_sigma_NO * c->helix_sigma_NO);
    } else if (strcmp(sec, "[Fragment_Memory]") == 0) {
      c->frag_mem_flag = 1;
      c->k_frag_mem = tk_double(&in);
      tk_str(&in, c->frag_mems_file, sizeof c->frag_mems_file);
      tk_str(&in, c->fm_gamma_file, sizeof c->fm_gamma_file);
    } else if (strcmp(sec, "[Epsilon]") == 0) {
      c->epsilon = tk_double(&in);
    } else if (strcmp(sec, "[Fragment_Memory_Table]") == 0) { /* fb.cpp:310-320 */
      c->frag_mem_tb_flag = 1;
      c->k_frag_mem = tk_double(&in);
      tk_str(&in, c->frag_mems_file, sizeof c->frag_mems_file);
      tk_str(&in, c->fm_gamma_file, sizeof c->fm_gamma_file);
      c->tb_rmin = tk_double(&in); c->tb_rmax = tk_double(&in); c->tb_dr = tk_double(&in);
      c->tb_size = (int)((c->tb_rmax - c->tb_rmin) / c->tb_dr) + 2;
      c->frag_table_well_width = tk_double(&in);
      c->fm_use_table_flag = tk_int(&in);
      c->fm_sigma_exp = tk_double(&in);
    } else if (strcmp(sec, "[AMH-Go]") == 0) { /* fb.cpp:297-303 */
      c->amh_go_flag = 1;
      c->k_amh_go = tk_double(&in);
      c->amh_go_p = tk_double(&in);
      c->amh_go_rc = tk_double(&in);
      c->frustration_censoring_flag = tk_int(&in);
    } else if (strcmp(sec, "[Contact_Restraints]") == 0) { /* fb.cpp:321-327 */
      c->cont_rest_flag = 1;
      c->k_cont_rest = tk_double(&in);
      c->cr_sigma = tk_double(&in);
      tk_str(&in, c->cr_file, sizeof c->cr_file);
      c->cr_sigma_sq_inv = pow(c->cr_sigma, -2);
    } else if (strcmp(sec, "[Vector_Fragment_Memory]") == 0) { /* fb.cpp:328-333 */
      c->vec_frag_mem_flag = 1;
      c->k_vec_frag_mem = tk_double(&in);
      c->vfm_sigma = tk_double(&in);
      c->vfm_sigma_sq = c->vfm_sigma * c->vfm_sigma;
    } else if (strcmp(sec, "[Solvent_Barrier]") == 0) { /* fb.cpp:334-345 */
      c->ssb_flag = 1;
      c->k_solventb1 = tk_double(&in);
      c->ssb_rmin1 = tk_double(&in); c->ssb_rmax1 = tk_double(&in);
      c->k_solventb2 = tk_double(&in);
      c->ssb_rmin2 = tk_double(&in); c->ssb_rmax2 = tk_double(&in);
      c->ssb_kappa = tk_double(&in);
      c->ssb_ij_sep = tk_int(&in);
      c->ssb_rad_cor = tk_int(&in);
      for (j = 0; j < 20; ++j) c->ssb_rshift[j] = tk_double(&in);
    } else if (strcmp(sec, "[Membrane]") == 0) { /* fb.cpp:346-358 */
      c->memb_flag = 1;
      c->k_overall_memb = tk_double(&in);
      c->k_bin = tk_double(&in);
      for (j = 0; j < 3; ++j) c->memb_xo[j] = tk_double(&in);
      c->memb_pore_type = tk_int(&in);
      c->memb_len = tk_double(&in);
      c->rho0_max = tk_double(&in);
      c->rho0_distor = tk_double(&in);
      for (int a = 0; a < 3; ++a)
        for (j = 0; j < 4; ++j) c->g_memb[a][j] = tk_double(&in);
    } else if (strcmp(sec, "[Phosphorylation]") == 0) { /* fb.cpp:412-421 */
      if (!c->water_flag) { free(text); set_err(err, errlen, "Cannot run phosphorylation without water potential"); return -1; }
      c->phosph_flag = 1;
      c->k_hypercharge = tk_double(&in);
      c->n_phosph_res = tk_int(&in);
      if (c->n_phosph_res > 20) { free(text); set_err(err, errlen, "Number of phosphorylated residues may not exceed 20"); return -1; }
      for (j = 0; j < c->n_phosph_res; ++j) c->phosph_res[j] = tk_int(&in);
    } else if (strcmp(sec, "[DebyeHuckel]") == 0) { /* fb.cpp:466-478 */
      c->huckel_flag = 1;
      c->k_PlusPlus = tk_double(&in); c->k_MinusMinus = tk_double(&in); c->k_PlusMinus = tk_double(&in);
      c->k_screening = tk_double(&in);
      c->screening_length = tk_double(&in);
      c->debye_huckel_min_sep = tk_int(&in);
    } else if (strcmp(sec, "[Tertiary_Frustratometer]") == 0) { /* fb.cpp:385-397 */
      c->tert_frust_flag = 1;
      c->tf_cutoff = tk_double(&in);
      c->tf_ndecoys = tk_int(&in);
      c->tf_freq = tk_int(&in);
      tk_str(&in, sbuf, sizeof sbuf);
      c->tf_done = 0;
      if (strcmp(sbuf, "configurational") == 0) c->tf_mode = 0;
      else if (strcmp(sbuf, "mutational") == 0) c->tf_mode = 1;
      else if (strcmp(sbuf, "singleresidue") == 0) c->tf_mode = 2;
      else {
        free(text);
        set_err(err, errlen, "Only \"configurational\", \"mutational\", \"singleresidue\" are acceptable modes for the Tertiary_Frustratometer.");
        return -1;
      }
    } else if (strcmp(sec, "[Nmer_Frustratometer]") == 0) { /* fb.cpp:398-412 */
      c->nmer_frust_flag = 1;
      c->nf_size = tk_int(&in);
      c->nf_cutoff = tk_double(&in);
      c->nf_contacts = tk_int(&in);
      c->nf_ndecoys = tk_int(&in);
      c->nf_freq = tk_int(&in);
      c->nf_min_thr = tk_double(&in); c->nf_high_thr = tk_double(&in); c->nf_neutral = tk_int(&in);
      c->nf_trap = tk_int(&in); c->nf_draw_trap = tk_int(&in); c->nf_trap_sigma = tk_double(&in); c->nf_ss_frac = tk_double(&in);
      tk_str(&in, sbuf, sizeof sbuf);
      if (strcmp(sbuf, "pairwise") == 0) c->nf_mode = 0;
      else if (strcmp(sbuf, "singlenmer") == 0) c->nf_mode = 1;
      else {
        free(text);
        set_err(err, errlen, "Only \"pairwise\", \"singlenmer\" are acceptable modes for the Nmer_Frustratometer.");
        return -1;
      }
    } else if (strcmp(sec, "[Fragment_Frustratometer]") == 0) { /* fb.cpp:359-384 */
      if (!c->frag_mem_flag && !c->frag_mem_tb_flag) {
        free(text);
        set_err(err, errlen, "Cannot run Fragment_Frustratometer without Fragment_Memory or Fragment_Memory_Table.");
        return -1;
      }
      c->frag_frust_flag = 1;
      c->ff_ww = 1.0;
      tk_str(&in, sbuf, sizeof sbuf);
      if (strcmp(sbuf, "shuffle") == 0) {
        c->ff_mode = 0;
        tk_str(&in, c->ff_file, sizeof c->ff_file);
        c->ff_ncalc = tk_int(&in);
        c->ff_freq = tk_int(&in);
      } else if (strcmp(sbuf, "read") == 0) {
        c->ff_mode = 1;
        tk_str(&in, c->ff_file, sizeof c->ff_file);
        c->ff_freq = tk_int(&in);
        c->ff_ww = tk_double(&in);
        c->ff_seqsep = tk_int(&in);
        c->ff_seqsep_gamma = tk_double(&in);
        c->ff_norm = tk_int(&in);
      } else {
        free(text);
        set_err(err, errlen, "Only \"shuffle\" and \"read\" are acceptable modes for the Fragment_Frustratometer.");
        return -1;
      }
    } else if (strcmp(sec, "[Amylometer]") == 0 || strcmp(sec, "[Selection_Temperature]") == 0 ||
               strcmp(sec, "[Monte_Carlo_Seq_Opt]") == 0 || strcmp(sec, "[Optimization]") == 0 ||
               strcmp(sec, "[Burial_Optimization]") == 0 ||
               strcmp(sec, "[DebyeHuckel_Optimization]") == 0) {
      free(text);
      set_err(err, errlen, "Section %s is enabled but is not part of this build's force path", sec);
      return -2;
    } else if (strcmp(sec, "[Shuffler]") == 0) {
      int fl = tk_int(&in);
      tk_str(&in, sbuf, sizeof sbuf);
      if (fl == 1) { free(text); set_err(err, errlen, "Section [Shuffler] is enabled but is not part of this build's force path"); return -2; }
    } else if (strcmp(sec, "[Mutate_Sequence]") == 0) {
      int fl = tk_int(&in);
      tk_str(&in, sbuf, sizeof sbuf);
      if (fl == 1) { free(text); set_err(err, errlen, "Section [Mutate_Sequence] is enabled but is not part of this build's force path"); return -2; }
    }
  }
  free(text);
  return 0;
}

static int read_table_file(const char *dir, const char *name, double *dst, int count, int stride, char *err, int errlen)
{
  (void)stride;
  char path[1024];
  join_path(path, sizeof path, dir, name);
  char *t = read_file(path, NULL);
  if (!t) { set_err(err, errlen, "File %s doesn't exist", name); return -1; }
  tokin in = {t, 0, 0};
  for (int k = 0; k < count; ++k) dst[k] = tk_double(&in);
  int bad = in.fail;
  free(t);
  if (bad) { set_err(err, errlen, "DSSP file format error"); return -1; }
  return 0;
}

static int fm_table_build(or_ctx *c, char *err, int errlen);

/* The C library's rand() as glibc implements it (the additive feedback
 * generator of its random_r, TYPE_3: degree 31, separation 3): srand(seed)
 * fills r[0] = seed, r[i] = 16807 r[i-1] mod (2^31 - 1) for i < 31, copies
 * r[31..33] = r[0..2], then r[i] = r[i-31] + r[i-3] (mod 2^32), and the
 * k-th rand() is r[344 + k] >> 1.  The reference draws its decoys with
 * rand() % n (fb.cpp:5550-5556) and never seeds it. */
void or_rng_seed(or_ctx *c, unsigned seed)
{
  int64_t r[34];
  r[0] = seed ? seed : 1;
  for (int i = 1; i < 31; ++i) r[i] = (16807 * r[i - 1]) % 2147483647;
  for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
  for (int i = 0; i < 34; ++i) c->rng[i] = (uint32_t)r[i];
  c->rng_k = 34;
  for (int i = 34; i < 344; ++i) or_rng_next(c);
}

int or_rng_next(or_ctx *c)
{
  /* ring of the last 34 values: r[k] at rng[k % 34] */
  int k = c->rng_k;
  uint32_t v = c->rng[(k - 31) % 34] + c->rng[(k - 3) % 34];
  c->rng[k % 34] = v;
  c->rng_k = k + 1;
  if (c->rng_k >= 34 * 1000000) c->rng_k = c->rng_k % 34 + 34;
  return (int)(v >> 1);
}

/* read_contact_restraints_file, fb.cpp:1106-1170: records "i1 i2 w r0"
 * (1-based residues) until one fails to read; w *= k_cont_rest (already
 * scaled by epsilon); per lower residue the entries sorted by the upper
 * one (qsort, fb.cpp:1163-1166; ties kept in file order here, the
 * reference's qsort leaves their order unspecified); the global cutoff
 * (r0_max + 4.29 sigma)^2 (fb.cpp:1147-1148). */
static int read_contact_restraints(or_ctx *c, const char *workdir, char *err, int errlen)
{
  char path[1024];
  join_path(path, sizeof path, workdir, c->cr_file);
  char *t = read_file(path, NULL);
  if (!t) { set_err(err, errlen, "File for Contact Restraints potential doesn't exist"); return -1; }
  tokin s = {t, 0, 0};
  int cap = 0, n = c->n;
  double r0_max = 0.0;
  int *i1 = NULL, *i2 = NULL;
  double *w = NULL, *r0 = NULL;
  c->n_cr = 0;
  for (;;) {
    int a = tk_int(&s), b = tk_int(&s);
    double ww = tk_double(&s), rr = tk_double(&s);
    if (s.fail) break;
    int j1 = (a < b ? a : b) - 1, j2 = (a < b ? b : a) - 1;
    if (j1 < 0 || j1 >= n || j2 < 0 || j2 >= n) {
      free(t); free(i1); free(i2); free(w); free(r0);
      set_err(err, errlen, "Indices out of range in Contact Restraints potential file");
      return -1;
    }
    if (c->n_cr == cap) {
      cap = cap ? 2 * cap : 64;
      i1 = (int *)realloc(i1, cap * sizeof(int)); i2 = (int *)realloc(i2, cap * sizeof(int));
      w = (double *)realloc(w, cap * sizeof(double)); r0 = (double *)realloc(r0, cap * sizeof(double));
    }
    i1[c->n_cr] = j1; i2[c->n_cr] = j2; w[c->n_cr] = ww * c->k_cont_rest; r0[c->n_cr] = rr;
    if (rr > r0_max) r0_max = rr;
    c->n_cr++;
  }
  free(t);
  /* counting sort by i1, then insertion sort by i2 within each (stable) */
  int m = c->n_cr;
  c->cr_start = (int *)calloc((size_t)n + 1, sizeof(int));
  for (int k = 0; k < m; ++k) c->cr_start[i1[k] + 1]++;
  for (int i = 0; i < n; ++i) c->cr_start[i + 1] += c->cr_start[i];
  int *fill = (int *)calloc((size_t)n, sizeof(int));
  c->cr_i1 = (int *)malloc((m ? m : 1) * sizeof(int)); c->cr_i2 = (int *)malloc((m ? m : 1) * sizeof(int));
  c->cr_w = (double *)malloc((m ? m : 1) * sizeof(double)); c->cr_r0 = (double *)malloc((m ? m : 1) * sizeof(double));
  for (int k = 0; k < m; ++k) {
    int q = c->cr_start[i1[k]] + fill[i1[k]]++;
    c->cr_i1[q] = i1[k]; c->cr_i2[q] = i2[k]; c->cr_w[q] = w[k]; c->cr_r0[q] = r0[k];
  }
  for (int i = 0; i < n; ++i)
    for (int q = c->cr_start[i] + 1; q < c->cr_start[i + 1]; ++q)
      for (int p = q; p > c->cr_start[i] && c->cr_i2[p - 1] > c->cr_i2[p]; --p) {
        int ti = c->cr_i2[p]; c->cr_i2[p] = c->cr_i2[p - 1]; c->cr_i2[p - 1] = ti;
        double tw = c->cr_w[p]; c->cr_w[p] = c->cr_w[p - 1]; c->cr_w[p - 1] = tw;
        double tr = c->cr_r0[p]; c->cr_r0[p] = c->cr_r0[p - 1]; c->cr_r0[p - 1] = tr;
      }
  free(fill); free(i1); free(i2); free(w); free(r0);
  c->cr_dr_cutoff = c->cr_sigma * 4.29;
  c->cr_glob_cutoff_sq = pow(r0_max + c->cr_dr_cutoff, 2);
  return 0;
}

/* AMH-Go set-up, fb.cpp:695-743: amh-go.gamma (Gamma_Array), amh-go.gro
 * (one memory over the whole system, pos 0, len n), the censored pairs of
 * frustration_censored_contacts.dat (flag 1, 1-based) or the native
 * distances of go_rnative{CACA,CBCB,CACB}.dat (flag 2, n x n each). */
static int amhgo_setup(or_ctx *c, const char *workdir, char *err, int errlen)
{
  char path[1024];
  int n = c->n;
  join_path(path, sizeof path, workdir, "amh-go.gamma");
  char e2[256];
  if (gm_load(&c->amhg, path, e2, sizeof e2)) {
    if (strstr(e2, "Cannot read")) set_err(err, errlen, "Cannot read file amh-go.gamma");
    else set_err(err, errlen, "AMH_Go: %s", strchr(e2, ':') ? strchr(e2, ':') + 2 : e2);
    return -1;
  }
  if (c->amhg.frag_resty) { set_err(err, errlen, "AMH-Go: Wrong call of getGamma() function"); return -1; }
  join_path(path, sizeof path, workdir, "amh-go.gro");
  gro_file g;
  gro_parse(path, &g);
  int e = mem_load(&c->amh_mem, 0, 0, n, 1.0, &g, 0);
  free(g.r);
  if (e == 1) { set_err(err, errlen, "Cannot read file amh-go.gro"); return -1; }
  if (e == 2) { set_err(err, errlen, "AMH_Go: Wrong atom count in memory structure file"); return -1; }
  if (e == 3) { set_err(err, errlen, "AMH_Go: Unknown residue"); return -1; }
  if (c->frustration_censoring_flag == 1) {
    c->censor = (unsigned char *)calloc((size_t)n * n, 1);
    join_path(path, sizeof path, workdir, "frustration_censored_contacts.dat");
    char *t = read_file(path, NULL);
    if (t) {
      tokin s = {t, 0, 0};
      for (;;) {
        int a = tk_int(&s), b = tk_int(&s);
        if (s.fail) break;
        if (a < 1 || a > n || b < 1 || b > n) { free(t); set_err(err, errlen, "AMH-Go: censored contact out of range"); return -1; }
        c->censor[(size_t)(a - 1) * n + (b - 1)] = 1;
      }
      free(t);
    }
  } else if (c->frustration_censoring_flag == 2) {
    const char *names[3] = {"go_rnativeCACA.dat", "go_rnativeCBCB.dat", "go_rnativeCACB.dat"};
    double **dst[3] = {&c->rn_caca, &c->rn_cbcb, &c->rn_cacb};
    for (int q = 0; q < 3; ++q) {
      join_path(path, sizeof path, workdir, names[q]);
      char *t = read_file(path, NULL);
      if (!t) { set_err(err, errlen, "Go native distance file can't be read"); return -1; }
      tokin s = {t, 0, 0};
      *dst[q] = (double *)calloc((size_t)n * n, sizeof(double));
      for (size_t k = 0; k < (size_t)n * n; ++k) (*dst[q])[k] = tk_double(&s);
      int bad = s.fail;
      free(t);
      if (bad) { set_err(err, errlen, "go_rnative*.dat file format error"); return -1; }
    }
  }
  c->go_f = (double *)calloc((size_t)2 * n * 3, sizeof(double));
  return 0;
}

/* native distance of the atom pair (i, ia) - (j, ja), ia/ja FM_CA or FM_CB
 * (fb.cpp:3838-3843, 3904-3909: the CA-CB matrix serves both mixed orders) */
static double amhgo_rnative(const or_ctx *c, int i, int ia, int j, int ja)
{
  if (c->frustration_censoring_flag == 2) {
    size_t k = (size_t)i * c->n + j;
    if (ia == FM_CA && ja == FM_CA) return c->rn_caca[k];
    if (ia == FM_CB && ja == FM_CB) return c->rn_cbcb[k];
    return c->rn_cacb[k];
  }
  return mem_rf(&c->amh_mem, i, ia, j, ja);
}

/* compute_amhgo_normalization, fb.cpp:3787-3852: a = sum over chains ich,
 * residues i, atoms ia (CA, and CB unless glycine) of, after each chain
 * jch, |normi|^p with normi the running sum (not reset per jch) of the
 * gammas of native contacts (rnative < rc, |i - j| >= minSep); a /= 8 N
 * with N the end of the last chain. */
static void amhgo_normalization(or_ctx *c)
{
  double norm = 0.0;
  int iresn = 0;
  for (int ich = 0; ich < c->nch; ++ich) {
    int ires0 = c->ch_pos[ich] - 1;
    iresn = c->ch_pos[ich] + c->ch_len[ich] - 1;
    for (int i = ires0; i < iresn; ++i) {
      int it = restype(c->se[i]);
      for (int ia = FM_CA; ia <= FM_CB - (c->se[i] == 'G' ? 1 : 0); ++ia) {
        double normi = 0.0;
        for (int jch = 0; jch < c->nch; ++jch) {
          int jres0 = c->ch_pos[jch] - 1, jresn = c->ch_pos[jch] + c->ch_len[jch] - 1;
          for (int j = jres0; j < jresn; ++j) {
            int jt = restype(c->se[j]);
            for (int ja = FM_CA; ja <= FM_CB - (c->se[j] == 'G' ? 1 : 0); ++ja) {
              if (abs(i - j) < c->amhg.i_sep[0]) continue;
              if (amhgo_rnative(c, i, ia, j, ja) < c->amh_go_rc) normi += gm_get2(&c->amhg, it, jt, i, j);
            }
          }
          norm += pow(fabs(normi), c->amh_go_p);
        }
      }
    }
  }
  c->amh_go_norm = norm / (8 * iresn);
}

int or_create(const char *coeff_path, const char *seq_path, const char *workdir,
              or_ctx **out, char *err, int errlen)
{
  *out = NULL;
  or_ctx *c = (or_ctx *)calloc(1, sizeof(or_ctx));
  c->epsilon = 1.0;
  /* fb.cpp:141-146 defaults */
  c->an = 0.4831806; c->bn = 0.7032820; c->cn = -0.1864262;
  c->ap = 0.4436538; c->bp = 0.2352006; c->cp = 0.3211455;
  c->ah = 0.8409657; c->bh = 0.8929599; c->ch = -0.7338894;
  c->fm_sigma_exp = 0.15;
  or_rng_seed(c, 1);  /* rand() of a process that never called srand */
  int rc = parse_coeff(c, coeff_path, err, errlen);
  if (rc) { or_destroy(c); return rc; }

  /* fb.cpp:502-516 */
  double e = c->epsilon;
  for (int j = 0; j < 3; ++j) c->k_chain[j] *= e;
  c->k_chi *= e; c->k_rama *= e; c->k_water *= e; c->k_burial *= e; c->k_helix *= e;
  c->k_dssp *= e; c->k_global_P_AP *= e;
  c->k_amh_go *= e; c->k_cont_rest *= e;
  for (int j = 0; j < c->n_rama_par; j++) c->w[j] *= c->k_rama;
  for (int j = 0; j < c->n_rama_p_par; j++) c->w[j + 6] *= c->k_rama;

  /* cWell ranges, sml.h:363-368; cP_AP ranges, sml.h:111-115 */
  for (int k = 0; k < c->n_wells; ++k) {
    c->w_rmin_theta[k] = c->well_r_min[k] - 8.0 * 2.302585 / c->water_kappa;
    c->w_rmax_theta[k] = c->well_r_max[k] + 8.0 * 2.302585 / c->water_kappa;
    c->w_rmin_theta_sq[k] = c->w_rmin_theta[k] * c->w_rmin_theta[k];
    c->w_rmax_theta_sq[k] = c->w_rmax_theta[k] * c->w_rmax_theta[k];
  }
  if (c->helix_flag) {
    c->h_rmin_theta = c->helix_well_r_min[0] - 8.0 * 2.302585 / c->helix_kappa;
    c->h_rmax_theta = c->helix_well_r_max[0] + 8.0 * 2.302585 / c->helix_kappa;
    c->h_rmin_theta_sq = c->h_rmin_theta * c->h_rmin_theta;
    c->h_rmax_theta_sq = c->h_rmax_theta * c->h_rmax_theta;
  }
  if (c->p_ap_flag) {
    double drmax = c->P_AP_cut + (double)((int)(8.0 * 2.302585 / c->P_AP_pref) + 1);
    double drmin = c->P_AP_cut - (double)((int)(8.0 * 2.302585 / c->P_AP_pref) + 1);
    c->pap_drmax_sq = drmax * drmax;
    c->pap_drmin_sq = drmin * drmin;
  }

  /* sequence file, fb.cpp:538-553 */
  char *st = read_file(seq_path, NULL);
  if (!st) { or_destroy(c); set_err(err, errlen, "Sequence file was not found"); return -1; }
  tokin in = {st, 0, 0};
  size_t secap = 1024, selen = 0;
  c->se = (char *)calloc(secap, 1);
  char buf[65536];
  while (!in.eof && !in.fail) {
    buf[0] = '\0';
    tk_str(&in, buf, sizeof buf);
    if (buf[0] == '#' || buf[0] == '\0') continue;
    if (c->nch >= MAXCH) { free(st); or_destroy(c); set_err(err, errlen, "Too many chains"); return -1; }
    size_t L = strlen(buf);
    for (size_t q = 0; q < L; ++q) {
      if (buf[q] < 'A' || buf[q] > 'Z') { free(st); or_destroy(c); set_err(err, errlen, "Sequence file: unknown residue letter"); return -1; }
    }
    c->ch_pos[c->nch] = (int)selen + 1;
    while (selen + L + 2 > secap) { secap *= 2; c->se = (char *)realloc(c->se, secap); }
    memcpy(c->se + selen, buf, L);
    selen += L;
    c->se[selen] = '\0';
    c->ch_len[c->nch] = (int)L;
    c->nch++;
  }
  free(st);
  c->n = (int)selen;
  if (c->n < 1) { or_destroy(c); set_err(err, errlen, "Empty sequence"); return -1; }
  int n = c->n;
  c->chain_no = (int *)calloc((size_t)n, sizeof(int));
  for (int ch = 0; ch < c->nch; ++ch)
    for (int k = 0; k < c->ch_len[ch]; ++k) c->chain_no[c->ch_pos[ch] - 1 + k] = ch + 1;
  for (int k = 0; k < 12; ++k) c->aps[k] = (double *)calloc((size_t)n, sizeof(double));

  char path[1024];
  if (c->dssp_hdrgn_flag) { /* fb.cpp:555-593 */
    double tmp[800];
    if (read_table_file(workdir, "anti_HB", tmp, 800, 0, err, errlen)) { or_destroy(c); return -1; }
    for (int i = 0; i < 20; ++i) for (int j = 0; j < 20; ++j) { c->m_anti_HB[i][j][0] = tmp[i * 20 + j]; c->m_anti_HB[i][j][1] = tmp[400 + i * 20 + j]; }
    if (read_table_file(workdir, "anti_NHB", tmp, 800, 0, err, errlen)) { or_destroy(c); return -1; }
    for (int i = 0; i < 20; ++i) for (int j = 0; j < 20; ++j) { c->m_anti_NHB[i][j][0] = tmp[i * 20 + j]; c->m_anti_NHB[i][j][1] = tmp[400 + i * 20 + j]; }
    if (read_table_file(workdir, "para_HB", tmp, 800, 0, err, errlen)) { or_destroy(c); return -1; }
    for (int i = 0; i < 20; ++i) for (int j = 0; j < 20; ++j) { c->m_para_HB[i][j][0] = tmp[i * 20 + j]; c->m_para_HB[i][j][1] = tmp[400 + i * 20 + j]; }
    if (read_table_file(workdir, "para_one", c->m_para_one, 20, 0, err, errlen)) { or_destroy(c); return -1; }
    if (read_table_file(workdir, "anti_one", c->m_anti_one, 20, 0, err, errlen)) { or_destroy(c); return -1; }
  }
  if (c->ssweight_flag) { /* fb.cpp:595-604 */
    join_path(path, sizeof path, workdir, "ssweight");
    char *t = read_file(path, NULL);
    if (!t) { or_destroy(c); set_err(err, errlen, "File ssweight doesn't exist"); return -1; }
    tokin s = {t, 0, 0};
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < 12; ++i) c->aps[i][j] = c->ssweight[i] ? tk_double(&s) : 0.0;
    free(t);
  }
  if (c->cont_rest_flag && read_contact_restraints(c, workdir, err, errlen)) { or_destroy(c); return -1; }
  if (c->memb_flag) { /* fb.cpp:611-620 */
    join_path(path, sizeof path, workdir, "zim");
    char *t = read_file(path, NULL);
    if (!t) { or_destroy(c); set_err(err, errlen, "File zim doesn't exist"); return -1; }
    tokin s = {t, 0, 0};
    c->z_res = (int *)calloc((size_t)n, sizeof(int));
    for (int i = 0; i < n; ++i) c->z_res[i] = tk_int(&s);
    int bad = s.fail;
    free(t);
    if (bad) { or_destroy(c); set_err(err, errlen, "Membrane potential parameter file format error"); return -1; }
    /* the reference leaves V unset for any other class (fb.cpp:6163-6172) */
    for (int i = 0; i < n; ++i)
      if (c->z_res[i] < 1 || c->z_res[i] > 3) { or_destroy(c); set_err(err, errlen, "Membrane potential: zim class outside 1-3"); return -1; }
  }
  if (c->water_flag) { /* fb.cpp:623-641 */
    join_path(path, sizeof path, workdir, "gamma.dat");
    char *t = read_file(path, NULL);
    if (!t) { or_destroy(c); set_err(err, errlen, "File gamma.dat doesn't exist"); return -1; }
    tokin s = {t, 0, 0};
    for (int w = 0; w < c->n_wells; ++w)
      for (int i = 0; i < 20; ++i)
        for (int j = i; j < 20; ++j) {
          double g0 = tk_double(&s), g1 = tk_double(&s);
          if (s.fail) { free(t); or_destroy(c); set_err(err, errlen, "Water potential gamma.dat file format error"); return -1; }
          c->water_gamma[w][i][j][0] = c->water_gamma[w][j][i][0] = g0 * c->k_water;
          c->water_gamma[w][i][j][1] = c->water_gamma[w][j][i][1] = g1 * c->k_water;
        }
    free(t);
  }
  if (c->phosph_flag) { /* fb.cpp:643-682 */
    for (int w = 0; w < c->n_wells; ++w)
      for (int i = 0; i < 20; ++i)
        for (int j = i; j < 20; ++j)
          for (int k = 0; k < 2; ++k)
            c->phosph_water_gamma[w][i][j][k] = c->phosph_water_gamma[w][j][i][k] = c->water_gamma[w][i][j][k];
    /* serine (15) takes the hypercharged glutamate (6) column, in this
       loop order (entries written earlier in the loop are read later) */
    for (int w = 0; w < c->n_wells; ++w)
      for (int i = 0; i < 20; ++i) {
        double f = four_letter_map[i] == 1 ? c->k_hypercharge
                 : (four_letter_map[i] == 2 || four_letter_map[i] == 3) ? pow(c->k_hypercharge, 2) : 1.0;
        for (int k = 0; k < 2; ++k)
          c->phosph_water_gamma[w][i][15][k] = c->phosph_water_gamma[w][15][i][k] = c->phosph_water_gamma[w][i][6][k] * f;
      }
    c->phosph_map = (int *)calloc((size_t)n, sizeof(int));
    for (int j = 0; j < c->n_phosph_res; ++j)
      if (c->phosph_res[j] != 0) {
        if (c->phosph_res[j] < 1 || c->phosph_res[j] > n) { or_destroy(c); set_err(err, errlen, "Phosphorylation: residue index out of range"); return -1; }
        c->phosph_map[c->phosph_res[j] - 1] = 1;
      }
  }
  if (c->burial_flag) { /* fb.cpp:684-693 */
    if (!c->water_flag) { or_destroy(c); set_err(err, errlen, "Cannot use burial without water potential"); return -1; }
    join_path(path, sizeof path, workdir, "burial_gamma.dat");
    char *t = read_file(path, NULL);
    if (!t) { or_destroy(c); set_err(err, errlen, "File burial_gamma.dat doesn't exist"); return -1; }
    tokin s = {t, 0, 0};
    for (int i = 0; i < 20; ++i) for (int k = 0; k < 3; ++k) c->burial_gamma[i][k] = tk_double(&s);
    int bad = s.fail;
    free(t);
    if (bad) { or_destroy(c); set_err(err, errlen, "Burial potential burial_gamma.dat file format error"); return -1; }
  }
  if (c->frag_mem_flag || c->frag_mem_tb_flag) { /* fb.cpp:749-791 */
    join_path(path, sizeof path, workdir, c->fm_gamma_file);
    if (gm_load(&c->fmg, path, err, errlen)) { or_destroy(c); return -1; }
    if (read_mems(c, workdir, c->frag_mems_file, &c->mems, &c->n_mems, c->vec_frag_mem_flag, err, errlen)) { or_destroy(c); return -1; }
    c->ilen_fm_map = (int *)calloc((size_t)n, sizeof(int));
    c->frag_mem_map = (int **)calloc((size_t)n, sizeof(int *));
    int min_sep = c->fmg.i_sep[0];
    for (int k = 0; k < c->n_mems; ++k) {
      int pos = c->mems[k].pos, len = c->mems[k].len;
      if (pos + len > n) { or_destroy(c); set_err(err, errlen, "Fragment_Memory: Incorrectly defined memory fragment"); return -1; }
      for (int i = pos; i < pos + len - min_sep; ++i) {
        c->ilen_fm_map[i]++;
        c->frag_mem_map[i] = (int *)realloc(c->frag_mem_map[i], (size_t)c->ilen_fm_map[i] * sizeof(int));
        c->frag_mem_map[i][c->ilen_fm_map[i] - 1] = k;
      }
    }
  }
  if (c->frag_frust_flag) { /* fb.cpp:794-855: the decoy library and its map */
    if (read_mems(c, workdir, c->ff_file, &c->ff_mems, &c->ff_nm, 0, err, errlen)) { or_destroy(c); return -1; }
    if (c->ff_mode == 1) c->ff_ncalc = c->ff_nm + 1;
    if (c->ff_ncalc < 2) { or_destroy(c); set_err(err, errlen, "Fragment_Frustratometer: fewer than one decoy calculation"); return -1; }
    c->ff_ilen = (int *)calloc((size_t)n, sizeof(int));
    c->ff_map = (int **)calloc((size_t)n, sizeof(int *));
    c->ff_mean = (double *)calloc((size_t)n, sizeof(double));
    c->ff_var = (double *)calloc((size_t)n, sizeof(double));
    if (ff_build_map(c)) { or_destroy(c); set_err(err, errlen, "Fragment_Frustratometer: Incorrectly defined memory fragment"); return -1; }
  }
  if (c->frag_mem_tb_flag && fm_table_build(c, err, errlen)) { or_destroy(c); return -1; }
  if (c->vec_frag_mem_flag && !(c->frag_mem_flag || c->frag_mem_tb_flag)) {
    /* the reference reads the memories only for those sections (fb.cpp:749-791) */
    or_destroy(c); set_err(err, errlen, "Vector_Fragment_Memory needs the memories of [Fragment_Memory]"); return -1;
  }
  if (c->amh_go_flag && amhgo_setup(c, workdir, err, errlen)) { or_destroy(c); return -1; }
  if (c->huckel_flag) { /* fb.cpp:1018-1046 */
    join_path(path, sizeof path, workdir, "charge_on_residues.dat");
    char *t = read_file(path, NULL);
    if (!t) { or_destroy(c); set_err(err, errlen, "File charge_on_residues.dat doesn't exist"); return -1; }
    tokin s = {t, 0, 0};
    c->charge = (double *)calloc((size_t)n, sizeof(double));
    int total = tk_int(&s);
    for (int k = 0; k < total && !s.fail; ++k) {
      int rn = tk_int(&s);
      double q = tk_double(&s);
      if (s.fail) break;
      if (rn < 1 || rn > n) { free(t); or_destroy(c); set_err(err, errlen, "charge_on_residues.dat: residue number out of range"); return -1; }
      c->charge[rn - 1] = q;
    }
    free(t);
  }
  /* fb.cpp:1048-1066 */
  double plc = 8.0;
  if (c->dssp_hdrgn_flag) plc = fmax(plc, c->dssp_hdrgn_cut);
  if (c->p_ap_flag) plc = fmax(plc, sqrt(c->pap_cutoff_sq));
  if (c->water_flag) for (int i = 0; i < c->n_wells; ++i) plc = fmax(plc, c->w_rmax_theta[i]);
  if (c->helix_flag) plc = fmax(plc, c->h_rmax_theta);
  if (c->helix_flag) plc = fmax(plc, calc_exp_helix_cutoff(c));
  if (c->helix_flag) plc = fmax(plc, c->helix_cutoff);
  if (c->cont_rest_flag) plc = fmax(plc, sqrt(c->cr_glob_cutoff_sq));
  if (c->huckel_flag) plc = fmax(plc, 2.0 * c->screening_length / c->k_screening);
  if (c->ssb_flag) {
    double cut = 0.0;
    if (c->ssb_rad_cor)
      for (int i = 0; i < 20; ++i) cut = fmax(cut, 2.0 * c->ssb_rshift[i]);
    cut += fmax(c->ssb_rmax1, c->ssb_rmax2) + 10.0 / c->ssb_kappa;
    plc = fmax(plc, cut);
  }
  c->pair_list_cutoff = plc;
  if (c->amh_go_flag) amhgo_normalization(c);

  /* excluded volume deck defaults (CoordinatesToWorkLammpsDataFile.py:314-322) */
  c->ev_p = 2;
  memset(c->ev_lambda, 0, sizeof c->ev_lambda);
  for (int a = 0; a < 6; ++a) for (int b = 0; b < 6; ++b) { c->ev_cut_short[a][b] = 3.5; c->ev_cut_long[a][b] = 3.5; }
  int tp[4][2] = {{1, 1}, {1, 4}, {4, 4}, {3, 3}};
  double cl[4] = {4.5, 4.5, 4.5, 3.5};
  for (int k = 0; k < 4; ++k) {
    int a = tp[k][0], b = tp[k][1];
    c->ev_lambda[a][b] = c->ev_lambda[b][a] = 20.0;
    c->ev_cut_short[a][b] = c->ev_cut_short[b][a] = 3.5;
    c->ev_cut_long[a][b] = c->ev_cut_long[b][a] = cl[k];
  }
  c->bond_k = 200.0;
  c->bond_r0[0] = 3.77; c->bond_r0[1] = 2.41; c->bond_r0[2] = 2.50; c->bond_r0[3] = 1.54; c->bond_r0[4] = 1.54;

  c->xn = (double *)calloc((size_t)n * 3, sizeof(double));
  c->xh = (double *)calloc((size_t)n * 3, sizeof(double));
  c->xcp = (double *)calloc((size_t)n * 3, sizeof(double));
  double **arrs[] = {&c->water_ro, &c->helix_ro, &c->w_sig, &c->w_sig_prd, &c->h_sig, &c->h_sig_prd,
                     &c->water_xi, &c->helix_xi_1, &c->helix_xi_2, &c->burial_force};
  for (size_t k = 0; k < sizeof arrs / sizeof arrs[0]; ++k) *arrs[k] = (double *)calloc((size_t)n, sizeof(double));
  *out = c;
  return 0;
}

void or_destroy(or_ctx *c)
{
  if (!c) return;
  for (int k = 0; k < c->n_mems; ++k) { free(c->mems[k].se); free(c->mems[k].rf0); free(c->mems[k].rf1); free(c->mems[k].vmf); }
  free(c->mems);
  for (int k = 0; k < c->ff_nm; ++k) { free(c->ff_mems[k].se); free(c->ff_mems[k].rf0); free(c->ff_mems[k].rf1); free(c->ff_mems[k].vmf); }
  free(c->ff_mems);
  if (c->ff_map) for (int i = 0; i < c->n; ++i) free(c->ff_map[i]);
  free(c->ff_map); free(c->ff_ilen); free(c->ff_mean); free(c->ff_var);
  if (c->frag_mem_map) for (int i = 0; i < c->n; ++i) free(c->frag_mem_map[i]);
  free(c->frag_mem_map); free(c->ilen_fm_map);
  free(c->fmg.gamma);
  if (c->tb_e) {
    size_t nt = (size_t)4 * c->n * c->tb_nbrs;
    for (size_t k = 0; k < nt; ++k) { free(c->tb_e[k]); free(c->tb_f[k]); }
    free(c->tb_e); free(c->tb_f);
  }
  for (int k = 0; k < 12; ++k) free(c->aps[k]);
  free(c->se); free(c->chain_no);
  free(c->xn); free(c->xh); free(c->xcp);
  free(c->water_ro); free(c->helix_ro); free(c->w_sig); free(c->w_sig_prd); free(c->h_sig); free(c->h_sig_prd);
  free(c->water_xi); free(c->helix_xi_1); free(c->helix_xi_2); free(c->burial_force);
  free(c->amhg.gamma); free(c->amh_mem.se); free(c->amh_mem.rf0); free(c->amh_mem.rf1); free(c->amh_mem.vmf);
  free(c->censor); free(c->rn_caca); free(c->rn_cbcb); free(c->rn_cacb); free(c->go_f);
  free(c->cr_i1); free(c->cr_i2); free(c->cr_start); free(c->cr_w); free(c->cr_r0);
  free(c->z_res); free(c->phosph_map); free(c->charge);
  free(c);
}

int or_n_residues(const or_ctx *c) { return c->n; }
int or_chain_of(const or_ctx *c, int i) { return c->chain_no[i] - 1; }
char or_seq_at(const or_ctx *c, int i) { return c->se[i]; }
double or_pair_list_cutoff(const or_ctx *c) { return c->pair_list_cutoff; }

void or_set_list_skin(or_ctx *c, double skin) { c->list_skin = skin; }

void or_set_bonds(or_ctx *c, double k, const double *r0)
{
  c->bond_k = k;
  for (int q = 0; q < 5; ++q) c->bond_r0[q] = r0[q];
}

void or_set_vexcluded(or_ctx *c, int p, const double *lam, const double *cs, const double *clg)
{
  c->ev_p = p;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) {
      c->ev_lambda[a][b] = lam[a * 6 + b];
      c->ev_cut_short[a][b] = cs[a * 6 + b];
      c->ev_cut_long[a][b] = clg[a * 6 + b];
    }
}

/* ------------------------------------------------------------------ */
/* virtual atoms, fb.cpp:7113-7138                                      */

static void build_virtual(or_ctx *c)
{
  int n = c->n;
  for (int i = 0; i < n; ++i) {
    int im1 = i > 0 ? i - 1 : -1;
    if (im1 != -1 && !is_first(c, i)) {
      for (int k = 0; k < 3; ++k) {
        XN(i)[k] = c->an * XCA(im1)[k] + c->bn * XCA(i)[k] + c->cn * XO(im1)[k];
        XH(i)[k] = c->ah * XCA(im1)[k] + c->bh * XCA(i)[k] + c->ch * XO(im1)[k];
        XCP(im1)[k] = c->ap * XCA(im1)[k] + c->bp * XCA(i)[k] + c->cp * XO(im1)[k];
      }
    } else {
      XN(i)[0] = XN(i)[1] = XN(i)[2] = 0.0;
      XH(i)[0] = XH(i)[1] = XH(i)[2] = 0.0;
      if (im1 != -1) XCP(im1)[0] = XCP(im1)[1] = XCP(im1)[2] = 0.0;
    }
  }
  XCP(n - 1)[0] = XCP(n - 1)[1] = XCP(n - 1)[2] = 0.0;
}

/* ------------------------------------------------------------------ */
/* local terms                                                          */

/* compute_chain_potential, fb.cpp:2062-2170 */
static void chain_term(or_ctx *c, int i)
{
  double dx[3], r, dr, force;
  double *E = &c->energy[OR_ET_CHAIN];
  if (!is_first(c, i) && c->se[i] != 'G') {
    int im1 = i - 1;
    for (int k = 0; k < 3; ++k) dx[k] = XN(i)[k] - XCB(i)[k];
    r = sqrt(dot3(dx, dx));
    dr = r - c->r_ncb0;
    force = 2 * c->k_chain[0] * dr / r;
    *E += c->k_chain[0] * dr * dr;
    fadd(FCA(im1), -c->an * force, dx);
    fadd(FO(im1), -c->cn * force, dx);
    fadd(FCA(i), -c->bn * force, dx);
    fadd(FCB(i), force, dx);
  }
  if (!is_last(c, i) && c->se[i] != 'G') {
    int ip1 = i + 1;
    for (int k = 0; k < 3; ++k) dx[k] = XCP(i)[k] - XCB(i)[k];
    r = sqrt(dot3(dx, dx));
    dr = r - c->r_cpcb0;
    force = 2 * c->k_chain[1] * dr / r;
    *E += c->k_chain[1] * dr * dr;
    fadd(FCA(ip1), -c->bp * force, dx);
    fadd(FCA(i), -c->ap * force, dx);
    fadd(FO(i), -c->cp * force, dx);
    fadd(FCB(i), force, dx);
  }
  if (!is_first(c, i) && !is_last(c, i)) {
    int im1 = i - 1, ip1 = i + 1;
    for (int k = 0; k < 3; ++k) dx[k] = XN(i)[k] - XCP(i)[k];
    r = sqrt(dot3(dx, dx));
    dr = r - c->r_ncp0;
    force = 2 * c->k_chain[2] * dr / r;
    *E += c->k_chain[2] * dr * dr;
    fadd(FCA(im1), -c->an * force, dx);
    fadd(FO(im1), -c->cn * force, dx);
    fadd(FCA(ip1), c->bp * force, dx);
    fadd(FCA(i), -(c->bn - c->ap) * force, dx);
    fadd(FO(i), c->cp * force, dx);
  }
}

/* compute_shake, fb.cpp:2172-2237 (k_shake is not pre-scaled: 2*epsilon*k) */
static void shake_term(or_ctx *c, int i)
{
  double dx[3], r, dr, force;
  double ks = c->epsilon * c->k_shake;
  double *E = &c->energy[OR_ET_SHAKE];
  if (!is_last(c, i)) {
    for (int k = 0; k < 3; ++k) dx[k] = XCA(i)[k] - XCA(i + 1)[k];
    r = sqrt(dot3(dx, dx)); dr = r - c->r_sh1; force = 2 * ks * dr / r;
    *E += ks * dr * dr;
    fadd(FCA(i), -force, dx); fadd(FCA(i + 1), force, dx);
  }
  for (int k = 0; k < 3; ++k) dx[k] = XCA(i)[k] - XO(i)[k];
  r = sqrt(dot3(dx, dx)); dr = r - c->r_sh2; force = 2 * ks * dr / r;
  *E += ks * dr * dr;
  fadd(FCA(i), -force, dx); fadd(FO(i), force, dx);
  if (!is_last(c, i)) {
    for (int k = 0; k < 3; ++k) dx[k] = XO(i)[k] - XCA(i + 1)[k];
    r = sqrt(dot3(dx, dx)); dr = r - c->r_sh3; force = 2 * ks * dr / r;
    *E += ks * dr * dr;
    fadd(FO(i), -force, dx); fadd(FCA(i + 1), force, dx);
  }
}

/* compute_chi_potential, fb.cpp:2239-2339 */
static void chi_term(or_ctx *c, int i)
{
  double a[3], b[3], cc[3], axb[3], cxa[3], bxc[3], aprl[3], bprl[3], cprl[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = XCP(i)[k] - XCA(i)[k];
    b[k] = XCA(i)[k] - XN(i)[k];
    cc[k] = XCA(i)[k] - XCB(i)[k];
  }
  double arvsq = 1.0 / dot3(a, a), brvsq = 1.0 / dot3(b, b), crvsq = 1.0 / dot3(cc, cc);
  double norm = sqrt(arvsq * brvsq * crvsq);
  cross3(a, b, axb); cross3(cc, a, cxa); cross3(b, cc, bxc);
  double chi = dot3(axb, cc) * norm;
  for (int k = 0; k < 3; ++k) {
    aprl[k] = norm * bxc[k] - arvsq * chi * a[k];
    bprl[k] = norm * cxa[k] - brvsq * chi * b[k];
    cprl[k] = norm * axb[k] - crvsq * chi * cc[k];
  }
  double dchi = chi - c->chi0;
  double force = 2 * c->k_chi * dchi;
  c->energy[OR_ET_CHI] += c->k_chi * dchi * dchi;
  if (!is_first(c, i)) {
    fadd(FCA(i - 1), c->an * force, bprl);
    fadd(FO(i - 1), c->cn * force, bprl);
  }
  if (!is_last(c, i)) fadd(FCA(i + 1), -c->bp * force, aprl);
  for (int k = 0; k < 3; ++k) FCA(i)[k] -= (cprl[k] + (1 - c->bn) * bprl[k] + (c->ap - 1) * aprl[k]) * force;
  fadd(FO(i), -c->cp * force, aprl);
  fadd(FCB(i), force, cprl);
}

/* calcDihedralAndSlopes, fb.cpp:2341-2453.  slopes[ang][atom][xyz],
 * atoms CA0=CA(i-1), CA1=CA(i), CA2=CA(i+1), O0=O(i-1), O1=O(i) */
static double dihedral_slopes(or_ctx *c, int i, int phi, double ys[5][3], double xs[5][3])
{
  double a[3], b[3], cc[3], bxa[3], cxa[3], cxb[3];
  double an = c->an, bn = c->bn, cn = c->cn, ap = c->ap, bp = c->bp, cp = c->cp;
  for (int k = 0; k < 3; ++k) {
    if (phi) {
      a[k] = XCP(i)[k] - XCA(i)[k];
      b[k] = XCA(i)[k] - XN(i)[k];
      cc[k] = XN(i)[k] - XCP(i - 1)[k];
    } else {
      a[k] = XN(i + 1)[k] - XCP(i)[k];
      b[k] = XCP(i)[k] - XCA(i)[k];
      cc[k] = XCA(i)[k] - XN(i)[k];
    }
  }
  double adb = dot3(a, b), bdc = dot3(b, cc), adc = dot3(a, cc), b2 = dot3(b, b), bm = sqrt(b2);
  cross3(b, a, bxa); cross3(cc, a, cxa); cross3(cc, b, cxb);
  double cdbxa = dot3(cc, bxa);
  double Y = bm * cdbxa, X = adb * bdc - b2 * adc;
  double angle = atan2(Y, X);
  double X2Y2 = 1.0 / (X * X + Y * Y);
  double dy = X * X2Y2, dxx = -Y * X2Y2;
  b2 *= dxx; adb *= dxx; adc *= dxx; bdc *= dxx;
  cdbxa *= dy / bm;
  bm *= dy;
  for (int l = 0; l < 3; l++) {
    double h1, h2, h3, h4;
    if (phi) {
      h1 = cxb[l] * bm; h2 = cxa[l] * bm; h3 = bxa[l] * bm; h4 = b[l] * cdbxa;
      ys[0][l] = -an * h4 + (an - ap) * h3 + an * h2;
      ys[1][l] = (1.0 - bn) * h4 + (bn - bp) * h3 + (ap - 1.0) * h1 - (1.0 - bn) * h2;
      ys[2][l] = bp * h1;
      ys[3][l] = -cn * h4 + (cn - cp) * h3 + cn * h2;
      ys[4][l] = cp * h1;
      h1 = b[l] * bdc - cc[l] * b2;
      h2 = a[l] * bdc - 2.0 * b[l] * adc + cc[l] * adb;
      h3 = b[l] * adb - a[l] * b2;
      xs[0][l] = -an * h2 + (an - ap) * h3;
      xs[1][l] = (ap - 1.0) * h1 + (1.0 - bn) * h2 + (bn - bp) * h3;
      xs[2][l] = bp * h1;
      xs[3][l] = -cn * h2 + (cn - cp) * h3;
      xs[4][l] = cp * h1;
    } else {
      h1 = bxa[l] * bm; h2 = cxb[l] * bm; h3 = cxa[l] * bm; h4 = b[l] * cdbxa;
      ys[0][l] = -an * h1;
      ys[1][l] = (ap - 1.0) * h4 + (1.0 - bn) * h1 + (an - ap) * h2 - (ap - 1) * h3;
      ys[2][l] = bp * h4 + (bn - bp) * h2 - bp * h3;
      ys[3][l] = -cn * h1;
      ys[4][l] = cp * h4 + (cn - cp) * h2 - cp * h3;
      h1 = b[l] * bdc - cc[l] * b2;
      h2 = a[l] * bdc - 2.0 * b[l] * adc + cc[l] * adb;
      h3 = b[l] * adb - a[l] * b2;
      xs[0][l] = -an * h3;
      xs[1][l] = (an - ap) * h1 + (ap - 1.0) * h2 + (1.0 - bn) * h3;
      xs[2][l] = (bn - bp) * h1 + bp * h2;
      xs[3][l] = -cn * h3;
      xs[4][l] = (cn - cp) * h1 + cp * h2;
    }
  }
  return angle;
}

/* compute_rama_potential / compute_rama_force, fb.cpp:2455-2546 */
static void rama_term(or_ctx *c, int i)
{
  double ys[2][5][3], xs[2][5][3];
  double phi = dihedral_slopes(c, i, 1, ys[0], xs[0]);
  double psi = dihedral_slopes(c, i, 0, ys[1], xs[1]);
  int js = 0, je = c->n_rama_par;
  if (c->se[i] == 'P' && c->rama_p_flag) { js = 6; je = 6 + c->n_rama_p_par; }
  double *atoms[5] = {FCA(i - 1), FCA(i), FCA(i + 1), FO(i - 1), FO(i)};
  for (int j = js; j < je; j++) {
    if (c->ssweight[j] && c->aps[j][i] == 0.0) continue;
    double cphi = cos(phi + c->phi0[j]) - 1.0, cpsi = cos(psi + c->psi0[j]) - 1.0;
    double pc = c->phiw[j] * cphi, sc = c->psiw[j] * cpsi;
    double V = c->w[j] * exp(-cphi * pc - cpsi * sc);
    if (c->ssweight[j]) V *= c->aps[j][i];
    double force = 2.0 * V;
    double f1[2] = {force * pc * sin(phi + c->phi0[j]), force * sc * sin(psi + c->psi0[j])};
    c->energy[OR_ET_RAMA] += -V;
    for (int ia = 0; ia < 2; ++ia)
      for (int at = 0; at < 5; ++at)
        for (int l = 0; l < 3; ++l) atoms[at][l] += f1[ia] * (ys[ia][at][l] + xs[ia][at][l]);
  }
}

/* ------------------------------------------------------------------ */
/* fragment memory, fb.cpp:4079-4177                                    */

static void fm_term(or_ctx *c, int i)
{
  int n = c->n;
  int itype = restype(c->se[i]);
  int maxsep = c->fmg.i_sep[c->fmg.nseq_cl], minsep = c->fmg.i_sep[0];
  const int ity[4] = {FM_CA, FM_CA, FM_CB, FM_CB}, jty[4] = {FM_CA, FM_CB, FM_CA, FM_CB};
  for (int ifm = 0; ifm < c->ilen_fm_map[i]; ++ifm) {
    const or_mem *m = &c->mems[c->frag_mem_map[i][ifm]];
    double ekw = c->epsilon * c->k_frag_mem * m->weight;
    int js = i + minsep, je = m->pos + m->len - 1;
    if (maxsep != -1 && je > i + maxsep) je = i + maxsep;
    if (je >= n) continue; /* reference aborts: "Missing residues in memory potential" */
    for (int j = js; j <= je; ++j) {
      int jtype = restype(c->se[j]);
      double sig2 = pow(abs(i - j), 2 * c->fm_sigma_exp);
      double g;
      if (!c->fmg.frag_resty) g = gm_get2(&c->fmg, itype, jtype, i, j);
      else g = gm_get4(&c->fmg, itype, jtype, restype(m->se[i - m->pos]), restype(m->se[j - m->pos]), i, j);
      double ekwg = ekw * g;
      const double *xi[4] = {XCA(i), XCA(i), XCB(i), XCB(i)}, *xj[4] = {XCA(j), XCB(j), XCA(j), XCB(j)};
      double *fi[4] = {FCA(i), FCA(i), FCB(i), FCB(i)}, *fj[4] = {FCA(j), FCB(j), FCA(j), FCB(j)};
      for (int k = 0; k < 4; ++k) {
        if (ity[k] == FM_CB && (c->se[i] == 'G' || m->se[i - m->pos] == 'G')) continue;
        if (jty[k] == FM_CB && (c->se[j] == 'G' || m->se[j - m->pos] == 'G')) continue;
        double dx[3] = {xi[k][0] - xj[k][0], xi[k][1] - xj[k][1], xi[k][2] - xj[k][2]};
        double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
        double rf = mem_rf(m, i, ity[k], j, jty[k]);
        double dr = r - rf, drsq = dr * dr;
        double V = -ekwg * exp(-drsq / (2.0 * sig2));
        c->energy[OR_ET_FRAGMEM] += V;
        double force = V * dr / (sig2 * r);
        fadd(fi[k], force, dx);
        fadd(fj[k], -force, dx);
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* fragment-memory table mode                                           */

/* FixBackbone::compute_fragment_memory_table, fb.cpp:4238-4323, with the
   allocation of fb.cpp:997-1017: for residue i, every memory covering it,
   j from i + minSep to the memory's end (capped at i + maxSep), and the
   four atom pairs (a CB skipped when the target or memory residue is a
   glycine), the curve of entry 4 tb_nbrs i + 4 (j - i - minSep) + k gets
   V(r) = -eps k_fm w gamma exp(-(r - rf)^2 / (2 sigma^2)) and
   V (r - rf) / (sigma^2 r) added at every grid point r = rmin + ir dr,
   sigma^2 = |i - j|^(2 fm_sigma_exp) ww^2 */
static int fm_table_build(or_ctx *c, char *err, int errlen)
{
  int n = c->n, minsep = c->fmg.i_sep[0], maxsep = c->fmg.i_sep[c->fmg.nseq_cl];
  c->tb_nbrs = maxsep != -1 ? maxsep - minsep + 1 : n - minsep;
  if (c->tb_nbrs <= 0 || c->tb_size <= 0) { set_err(err, errlen, "Fragment_Memory_Table: empty table"); return -1; }
  size_t nt = (size_t)4 * n * c->tb_nbrs;
  c->tb_e = (double **)calloc(nt, sizeof(double *));
  c->tb_f = (double **)calloc(nt, sizeof(double *));
  const int ity[4] = {FM_CA, FM_CA, FM_CB, FM_CB}, jty[4] = {FM_CA, FM_CB, FM_CA, FM_CB};
  for (int i = 0; i < n; ++i) {
    int itype = restype(c->se[i]);
    for (int ifm = 0; ifm < c->ilen_fm_map[i]; ++ifm) {
      const or_mem *m = &c->mems[c->frag_mem_map[i][ifm]];
      double ekw = c->epsilon * c->k_frag_mem * m->weight;
      int js = i + minsep, je = m->pos + m->len - 1;
      if (maxsep != -1 && je > i + maxsep) je = i + maxsep;
      if (je >= n) { set_err(err, errlen, "Missing residues in memory potential"); return -1; }
      for (int j = js; j <= je; ++j) {
        int jtype = restype(c->se[j]);
        double sig2 = pow(abs(i - j), 2 * c->fm_sigma_exp);
        sig2 = sig2 * c->frag_table_well_width * c->frag_table_well_width;
        double g;
        if (!c->fmg.frag_resty) g = gm_get2(&c->fmg, itype, jtype, i, j);
        else g = gm_get4(&c->fmg, itype, jtype, restype(m->se[i - m->pos]), restype(m->se[j - m->pos]), i, j);
        double ekwg = ekw * g;
        for (int k = 0; k < 4; ++k) {
          if (ity[k] == FM_CB && (c->se[i] == 'G' || m->se[i - m->pos] == 'G')) continue;
          if (jty[k] == FM_CB && (c->se[j] == 'G' || m->se[j - m->pos] == 'G')) continue;
          size_t itb = (size_t)4 * c->tb_nbrs * i + 4 * (j - js) + k;
          if (!c->tb_e[itb]) {
            c->tb_e[itb] = (double *)calloc((size_t)c->tb_size, sizeof(double));
            c->tb_f[itb] = (double *)calloc((size_t)c->tb_size, sizeof(double));
          }
          double rf = mem_rf(m, i, ity[k], j, jty[k]);
          for (int ir = 0; ir < c->tb_size; ++ir) {
            double r = c->tb_rmin + ir * c->tb_dr;
            double dr = r - rf, drsq = dr * dr;
            double V = -ekwg * exp(-drsq / (2 * sig2));
            c->tb_e[itb][ir] += V;
            c->tb_f[itb][ir] += V * dr / (sig2 * r);
          }
        }
      }
    }
  }
  return 0;
}

/* FixBackbone::table_fragment_memory, fb.cpp:4325-4429: linear
   interpolation of both curves at r, force f_i += ff dx, f_j -= ff dx.  A
   missing curve ends the pair (the reference returns, also for the atom
   pairs after it); a distance outside [rmin, rmax] is an error. */
static int fm_table_term(or_ctx *c, int i, int j)
{
  int minsep = c->fmg.i_sep[0], maxsep = c->fmg.i_sep[c->fmg.nseq_cl];
  if (j - i < minsep) return 0;
  if (maxsep != -1 && j - i > maxsep) return 0;
  int tb_j = j - i - minsep;
  size_t itb0 = (size_t)4 * c->tb_nbrs * i + 4 * tb_j;
  if (!c->tb_e[itb0]) return 0;
  const int ity[4] = {FM_CA, FM_CA, FM_CB, FM_CB}, jty[4] = {FM_CA, FM_CB, FM_CA, FM_CB};
  const double *xi[4] = {XCA(i), XCA(i), XCB(i), XCB(i)}, *xj[4] = {XCA(j), XCB(j), XCA(j), XCB(j)};
  double *fi[4] = {FCA(i), FCA(i), FCB(i), FCB(i)}, *fj[4] = {FCA(j), FCB(j), FCA(j), FCB(j)};
  for (int k = 0; k < 4; ++k) {
    if (c->se[i] == 'G' && ity[k] == FM_CB) continue;
    if (c->se[j] == 'G' && jty[k] == FM_CB) continue;
    double dx[3] = {xi[k][0] - xj[k][0], xi[k][1] - xj[k][1], xi[k][2] - xj[k][2]};
    double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    if (!(r >= c->tb_rmin && r <= c->tb_rmax)) {
      snprintf(c->last_err, sizeof c->last_err, "Table Fragment Memory: r is out of computed range.");
      return -1;
    }
    int ir = (int)((r - c->tb_rmin) / c->tb_dr);
    size_t itb = itb0 + k;
    if (!c->tb_e[itb]) return 0;
    if (ir < 0 || ir >= c->tb_size) {
      snprintf(c->last_err, sizeof c->last_err, "Table Fragment Memory: ir is out of range.");
      return -1;
    }
    double r1 = c->tb_rmin + (double)ir * c->tb_dr, r2 = c->tb_rmin + (double)(ir + 1) * c->tb_dr;
    double v1 = c->tb_e[itb][ir], v2 = c->tb_e[itb][ir + 1];
    double V = ((v2 - v1) * r + v1 * r2 - v2 * r1) / (r2 - r1);
    double f1 = c->tb_f[itb][ir], f2 = c->tb_f[itb][ir + 1];
    double ff = ((f2 - f1) * r + f1 * r2 - f2 * r1) / (r2 - r1);
    c->energy[OR_ET_FRAGMEM] += V;
    fadd(fi[k], ff, dx);
    fadd(fj[k], -ff, dx);
  }
  return 0;
}

/* the pair path's loop (fb.cpp:7513-7527): j from i + minSep to
   ch_pos + ch_len - 2, the last residue of i's chain (ch_pos is 1-based,
   fb.cpp:546), capped at i + maxSep */
static int fm_table_terms(or_ctx *c, int i)
{
  int minsep = c->fmg.i_sep[0], maxsep = c->fmg.i_sep[c->fmg.nseq_cl];
  int ch = c->chain_no[i] - 1;
  int jr0 = i + minsep, jrn = c->ch_pos[ch] + c->ch_len[ch] - 2;
  if (maxsep != -1 && i + maxsep < jrn) jrn = i + maxsep;
  for (int j = jr0; j <= jrn; ++j)
    if (fm_table_term(c, i, j)) return -1;
  return 0;
}

/* ------------------------------------------------------------------ */
/* pair machinery                                                       */

/* cWell::compute_theta_pair, sml.h:594-610 */
static void theta_pair(double kappa, double rmin, double rmax, double rmin_th, double rmax_th, double r,
                       double *th, double *prd)
{
  if (r < rmin_th || r > rmax_th) { *th = 0.0; *prd = 0.0; return; }
  double tmin = tanh(kappa * (r - rmin)), tmax = tanh(kappa * (rmax - r));
  *th = 0.25 * (1.0 + tmin) * (1.0 + tmax);
  *prd = kappa * (*th) * (tmax - tmin) / r;
}

static inline void wtheta(const or_ctx *c, int w, double r, double *th, double *prd)
{
  theta_pair(c->water_kappa, c->well_r_min[w], c->well_r_max[w], c->w_rmin_theta[w], c->w_rmax_theta[w], r, th, prd);
}
static inline void htheta(const or_ctx *c, double r, double *th, double *prd)
{
  theta_pair(c->helix_kappa, c->helix_well_r_min[0], c->helix_well_r_max[0], c->h_rmin_theta, c->h_rmax_theta, r, th, prd);
}

void or_theta_pair(const or_ctx *c, int helix, int well, double r, double *theta, double *prd_theta)
{
  if (helix) htheta(c, r, theta, prd_theta);
  else wtheta(c, well, r, theta, prd_theta);
}

/* density site: CB, or CA for Gly (fb.cpp:7738) */
static inline const double *site(const or_ctx *c, int i) { return c->se[i] == 'G' ? XCA(i) : XCB(i); }
static inline double *fsite(or_ctx *c, int i) { return c->se[i] == 'G' ? FCA(i) : FCB(i); }

static inline double rsq_of(const double *a, const double *b)
{
  double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  return d0 * d0 + d1 * d1 + d2 * d2;
}

/* cP_AP::compute, sml.h:176-203 */
static void pap_nu_pair(const or_ctx *c, int i, int j, double *nu, double *prd)
{
  double d[3] = {XCA(i)[0] - XCA(j)[0], XCA(i)[1] - XCA(j)[1], XCA(i)[2] - XCA(j)[2]};
  double drsq = pow(d[0], 2) + pow(d[1], 2) + pow(d[2], 2);
  or_pap_nu(c, drsq, nu, prd);
}

void or_pap_nu(const or_ctx *c, double drsq, double *nu, double *prd)
{
  if (drsq > c->pap_drmax_sq) { *nu = 0.0; *prd = 0.0; }
  else if (drsq < c->pap_drmin_sq) { *nu = 1.0; *prd = 0.0; }
  else {
    double dr = sqrt(drsq);
    double th = tanh(c->P_AP_pref * (c->P_AP_cut - dr));
    *nu = 0.5 * (1 + th);
    *prd = 0.5 * c->P_AP_pref * (1 - pow(th, 2)) / dr;
  }
}

/* cR::rNO / rHO, sml.h:273-301 */
static inline double rNO(const or_ctx *c, int i, int j)
{
  return sqrt(pow(XO(i)[0] - XN(j)[0], 2) + pow(XO(i)[1] - XN(j)[1], 2) + pow(XO(i)[2] - XN(j)[2], 2));
}
static inline double rHO(const or_ctx *c, int i, int j)
{
  return sqrt(pow(XO(i)[0] - XH(j)[0], 2) + pow(XO(i)[1] - XH(j)[1], 2) + pow(XO(i)[2] - XH(j)[2], 2));
}

/* DSSP hydrogen bond, compute_dssp_hdrgn fb.cpp:2902-3277 */
static void dssp_term(or_ctx *c, int i, int j)
{
  if (rNO(c, i, j) > c->dssp_hdrgn_cut) return;
  int i_rep = 1, i_AP = 1, i_P = 1, itheta[4];
  double lambda[4] = {0, 0, 0, 0}, R_NO[4], R_HO[4], theta[4], nu[2], prdnu[2], prd_theta[4][2], V[4];
  double dxnu[2][3], xNO[4][3], xHO[4][3], th;
  double tsa_HB[2], tsa_NHB[2], tsp_HB[2];
  int hb_class = 0;
  const char *se = c->se;
  int ic = c->chain_no[i] - 1, jc = c->chain_no[j] - 1;
  int i_ch_start = c->ch_pos[ic], j_ch_start = c->ch_pos[jc];
  int i_ch_end = c->ch_pos[ic] + c->ch_len[ic] - 1, j_ch_end = c->ch_pos[jc] + c->ch_len[jc] - 1;
  if (is_last(c, j) || se[j + 1] == 'P') i_rep = 0;
  if (is_first(c, i) || is_last(c, j) || se[i] == 'P') i_AP = 0;
  if (i >= i_ch_end - 2 || is_last(c, j) || se[i + 2] == 'P') i_P = 0;
#define ANTI_HB(a, b, k) c->m_anti_HB[restype(a)][restype(b)][k]
#define ANTI_NHB(a, b, k) c->m_anti_NHB[restype(a)][restype(b)][k]
#define PARA_HB(a, b, k) c->m_para_HB[restype(a)][restype(b)][k]
#define ANTI_ONE(a) c->m_anti_one[restype(a)]
#define PARA_ONE(a) c->m_para_one[restype(a)]
  for (int k = 0; k < 2; ++k) {
    if (i_AP) {
      tsa_HB[k] = 0.5 * ANTI_HB(se[i], se[j], k);
      tsa_NHB[k] = 0.25 * (ANTI_NHB(se[i + 1], se[j - 1], k) + ANTI_NHB(se[i - 1], se[j + 1], k));
    } else tsa_HB[k] = tsa_NHB[k] = 0.0;
    tsp_HB[k] = i_P ? PARA_HB(se[i + 1], se[j], k) : 0.0;
  }
  const double (*h)[9] = c->hbscl;
  int nr = c->n_rama_par;
  if (ic == jc && abs(j - i) < 45) {
    if (abs(j - i) < 4) {
      lambda[0] = -h[0][0]; lambda[1] = -h[0][1]; lambda[2] = 0.0; lambda[3] = 0.0;
      hb_class = 1;
    } else if (abs(j - i) < 18) {
      if (nr > 0 && c->aps[nr - 1][i] == 1.0 && c->aps[nr - 1][j] == 1.0) {
        lambda[0] = -h[1][0];
        lambda[1] = -h[1][1];
        lambda[2] = -h[1][2] - h[1][3] * tsa_HB[0] - h[1][4] * tsa_NHB[0] - h[1][5] * (ANTI_ONE(se[i]) + ANTI_ONE(se[j]));
        lambda[3] = -h[1][6];
      } else {
        lambda[0] = 0.0; lambda[1] = -h[1][1]; lambda[2] = 0.0; lambda[3] = 0.0;
      }
      hb_class = 2;
    } else {
      lambda[0] = -h[2][0];
      lambda[1] = -h[2][1];
      lambda[2] = -h[2][2] - h[2][3] * tsa_HB[1] - h[2][4] * tsa_NHB[1] - h[2][5] * (ANTI_ONE(se[i]) + ANTI_ONE(se[j]));
      lambda[3] = -h[2][6] - h[2][7] * tsp_HB[1] - h[2][8] * (PARA_ONE(se[i + 1]) + PARA_ONE(se[j]));
      hb_class = 3;
    }
  } else {
    lambda[0] = -h[3][0];
    lambda[1] = -h[3][1];
    lambda[2] = -h[3][2] - h[3][3] * tsa_HB[1] - h[3][4] * tsa_NHB[1] - h[3][5] * (ANTI_ONE(se[i]) + ANTI_ONE(se[j]));
    lambda[3] = -h[3][6] - h[3][7] * tsp_HB[1] - h[3][8] * (PARA_ONE(se[i + 1]) + PARA_ONE(se[j]));
    hb_class = 4;
  }
  nu[0] = prdnu[0] = 0.0;
  int nu0_on = (i - 2 >= i_ch_start - 1 && i + 2 < i_ch_end && hb_class != 2);
  if (nu0_on) {
    for (int k = 0; k < 3; ++k) dxnu[0][k] = XCA(i + 2)[k] - XCA(i - 2)[k];
    double rr = dot3(dxnu[0], dxnu[0]);
    if (rr > c->dssp_nu_cut1_sq) {
      rr = sqrt(rr);
      th = tanh(c->pref[0] * (rr - c->d_nu0));
      nu[0] = 0.5 * (1.0 + th);
      prdnu[0] = c->pref[0] * nu[0] * (1.0 - th) / rr;
    }
  } else nu[0] = 1.0;
  nu[1] = prdnu[1] = 0.0;
  int nu1_on = (j - 2 >= j_ch_start - 1 && j + 2 < j_ch_end && hb_class != 2);
  if (nu1_on) {
    for (int k = 0; k < 3; ++k) dxnu[1][k] = XCA(j + 2)[k] - XCA(j - 2)[k];
    double rr = dot3(dxnu[1], dxnu[1]);
    if (rr > c->dssp_nu_cut2_sq) {
      rr = sqrt(rr);
      th = tanh(c->pref[1] * (rr - c->d_nu0));
      nu[1] = 0.5 * (1.0 + th);
      prdnu[1] = c->pref[1] * nu[1] * (1.0 - th) / rr;
    }
  } else nu[1] = 1.0;
  if (nu0_on && nu[0] > 0.0) c->branch[OR_BR_DSSP_NU0]++;
  if (nu1_on && nu[1] > 0.0) c->branch[OR_BR_DSSP_NU1]++;
  if (nu[0] < DSSP_NU_DELTA || nu[1] < DSSP_NU_DELTA) { c->branch[OR_BR_DSSP_NU_EXIT]++; return; }
  c->branch[OR_BR_DSSP_CLASS1 + hb_class - 1]++;
  if (hb_class == 2 && lambda[0] != 0.0) c->branch[OR_BR_DSSP_BETA]++;

  itheta[0] = 1; itheta[1] = i_rep; itheta[2] = i_AP; itheta[3] = i_P;
  /* geometry of the four H-bond channels: (acceptor O, donor N/H) */
  int acc[4] = {i, i, j, j}, don[4] = {j, j + 1, i, i + 2};
  for (int k = 0; k < 4; ++k) {
    if (!itheta[k]) continue;
    R_NO[k] = rNO(c, acc[k], don[k]);
    R_HO[k] = rHO(c, acc[k], don[k]);
    for (int l = 0; l < 3; ++l) {
      xNO[k][l] = XO(acc[k])[l] - XN(don[k])[l];
      xHO[k][l] = XO(acc[k])[l] - XH(don[k])[l];
    }
  }
  for (int k = 0; k < 4; ++k) {
    if (itheta[k]) {
      double dNO = (R_NO[k] - c->NO_zero) * c->sigma_NO_sqinv;
      double dHO = (R_HO[k] - c->HO_zero) * c->sigma_HO_sqinv;
      theta[k] = exp(-0.5 * ((R_NO[k] - c->NO_zero) * dNO + (R_HO[k] - c->HO_zero) * dHO));
      prd_theta[k][0] = -dNO / R_NO[k];
      prd_theta[k][1] = -dHO / R_HO[k];
    } else {
      theta[k] = 0.0;
      prd_theta[k][0] = prd_theta[k][1] = 0.0;
    }
  }
  for (int k = 1; k < 4; ++k)
    if (itheta[k] && theta[k] > 1e-3) c->branch[OR_BR_DSSP_REP + k - 1]++;
  double ff1 = c->k_dssp * theta[0];
  V[0] = lambda[0];
  V[1] = lambda[1] * theta[1];
  V[2] = lambda[2] * theta[2];
  V[3] = lambda[3] * theta[3];
  double theta_sum = ff1 * (V[0] + V[1] + V[2] + V[3]);
  ff1 *= nu[0] * nu[1];
  for (int k = 0; k < 4; ++k) V[k] *= ff1;
  double VT = V[0] + V[1] + V[2] + V[3];
  if (ON(OR_ET_DSSP)) c->energy[OR_ET_DSSP] += VT;
  if (!ON(OR_ET_DSSP)) return;
  if (nu0_on) {
    double f = theta_sum * prdnu[0] * nu[1];
    fadd(FCA(i - 2), f, dxnu[0]);
    fadd(FCA(i + 2), -f, dxnu[0]);
  }
  if (nu1_on) {
    double f = theta_sum * nu[0] * prdnu[1];
    fadd(FCA(j - 2), f, dxnu[1]);
    fadd(FCA(j + 2), -f, dxnu[1]);
  }
  double Vk[4] = {VT, V[1], V[2], V[3]};
  for (int k = 0; k < 4; ++k) {
    if (!itheta[k]) continue;
    double p0 = prd_theta[k][0] * Vk[k], p1 = prd_theta[k][1] * Vk[k];
    int a = acc[k], d = don[k];
    for (int l = 0; l < 3; ++l) {
      FO(a)[l] -= p0 * xNO[k][l] + p1 * xHO[k][l];
      FCA(d - 1)[l] += c->an * p0 * xNO[k][l] + c->ah * p1 * xHO[k][l];
      FCA(d)[l] += c->bn * p0 * xNO[k][l] + c->bh * p1 * xHO[k][l];
      FO(d - 1)[l] += c->cn * p0 * xNO[k][l] + c->ch * p1 * xHO[k][l];
    }
  }
}

/* P_AP, compute_P_AP_potential fb.cpp:3279-3383 (i<j by residue index) */
static void pap_term(or_ctx *c, int i, int j)
{
  double nu_ij, prd_ij;
  pap_nu_pair(c, i, j, &nu_ij, &prd_ij);
  if (nu_ij < PAP_DELTA) return;
  int n = c->n, d = c->i_diff_P_AP;
  int ic = c->chain_no[i] - 1, jc = c->chain_no[j] - 1;
  int i_ch_end = c->ch_pos[ic] + c->ch_len[ic] - 1, j_ch_end = c->ch_pos[jc] + c->ch_len[jc] - 1;
  int mn = c->i_med_min, mx = c->i_med_max;
  int jmed_hi = (i + mx + 2 * d) < (i_ch_end - 1) ? (i + mx + 2 * d) : (i_ch_end - 1);
  int i_AP_med = ic == jc && i < i_ch_end - (mn + 2 * d) && j >= i + (mn + 2 * d) && j <= jmed_hi;
  int i_AP_long = (ic == jc && i < i_ch_end - (mx + 2 * d + 1) && j >= i + (mx + 2 * d + 1) && j < j_ch_end) ||
                  (ic != jc && i + d < n && j - d >= 0 && (c->chain_no[i + d] - 1) == ic && (c->chain_no[j - d] - 1) == jc);
  int i_P = (ic == jc && i < i_ch_end - (mx + 1 + d) && j >= i + (mx + 1) && j < i_ch_end - d) ||
            (ic != jc && i + d < n && j + d < n && (c->chain_no[i + d] - 1) == ic && (c->chain_no[j + d] - 1) == jc);
  int nr = c->n_rama_par;
  int beta = nr > 0 && c->aps[nr - 1][i] == 1.0 && c->aps[nr - 1][j] == 1.0;
  if (i_AP_med) c->branch[OR_BR_PAP_MED]++;
  if (i_AP_long) c->branch[ic == jc ? OR_BR_PAP_LONG : OR_BR_PAP_LONG_X]++;
  if (i_P) c->branch[ic == jc ? OR_BR_PAP_PAR : OR_BR_PAP_PAR_X]++;
  if (beta && (i_AP_med || i_AP_long || i_P)) c->branch[OR_BR_PAP_BETA]++;
  if (!ON(OR_ET_PAP)) return;
  if (i_AP_med || i_AP_long) {
    double K = beta ? (i_AP_med ? c->k_P_AP[0] : 0.0) + (i_AP_long ? c->k_P_AP[1] * c->k_betapred_P_AP : 0.0)
                    : (i_AP_med ? c->k_P_AP[0] : 0.0) + (i_AP_long ? c->k_P_AP[1] : 0.0);
    double nu2, prd2;
    pap_nu_pair(c, i + d, j - d, &nu2, &prd2);
    c->energy[OR_ET_PAP] += -c->k_global_P_AP * K * nu_ij * nu2;
    double dx0[3], dx1[3];
    for (int l = 0; l < 3; ++l) { dx0[l] = XCA(i)[l] - XCA(j)[l]; dx1[l] = XCA(i + d)[l] - XCA(j - d)[l]; }
    double f0 = c->k_global_P_AP * K * prd_ij * nu2, f1 = c->k_global_P_AP * K * nu_ij * prd2;
    fadd(FCA(i), -f0, dx0); fadd(FCA(j), f0, dx0);
    fadd(FCA(i + d), -f1, dx1); fadd(FCA(j - d), f1, dx1);
  }
  if (i_P) {
    double K = beta ? c->k_P_AP[2] * c->k_betapred_P_AP : c->k_P_AP[2];
    double nu2, prd2;
    pap_nu_pair(c, i + d, j + d, &nu2, &prd2);
    c->energy[OR_ET_PAP] += -c->k_global_P_AP * K * nu_ij * nu2;
    double dx0[3], dx1[3];
    for (int l = 0; l < 3; ++l) { dx0[l] = XCA(i)[l] - XCA(j)[l]; dx1[l] = XCA(i + d)[l] - XCA(j + d)[l]; }
    double f0 = c->k_global_P_AP * K * prd_ij * nu2, f1 = c->k_global_P_AP * K * nu_ij * prd2;
    fadd(FCA(i), -f0, dx0); fadd(FCA(j), f0, dx0);
    fadd(FCA(i + d), -f1, dx1); fadd(FCA(j + d), f1, dx1);
  }
}

/* compute_helix_dtheta_pair, fb.cpp:3706-3785 */
static void helix_term(or_ctx *c, int i, int j)
{
  if (rNO(c, i, j) > c->helix_cutoff) return;
  if (c->chain_no[i] != c->chain_no[j]) return;
  int it = restype(c->se[i]), jt = restype(c->se[j]);
  double R_NO = rNO(c, i, j), R_HO = rHO(c, i, j), xNO[3], xHO[3];
  for (int l = 0; l < 3; ++l) { xNO[l] = XO(i)[l] - XN(j)[l]; xHO[l] = XO(i)[l] - XH(j)[l]; }
  double h4i = c->h4prob[it];
  if (c->se[i] == 'P' && c->pro_accepter_flag) h4i = c->h4prob_pro_accepter;
  double prob_sum = h4i + c->h4prob[jt];
  c->branch[OR_BR_HELIX]++;
  double dNO = R_NO - c->helix_NO_zero, dHO = R_HO - c->helix_HO_zero;
  double dNOs = dNO * c->helix_sigma_NO_sqinv, dHOs = dHO * c->helix_sigma_HO_sqinv;
  double pt = -c->k_helix * prob_sum * exp(-0.5 * (dNO * dNOs + dHO * dHOs));
  double p0 = -dNOs / R_NO, p1 = -dHOs / R_HO;
  double gdiff = c->helix_gamma_w - c->helix_gamma_p;
  double sg = c->helix_gamma_p + gdiff * c->h_sig[i] * c->h_sig[j];
  c->helix_xi_1[i] = gdiff * pt * c->h_sig_prd[i] * c->h_sig[j];
  c->helix_xi_2[i] = gdiff * pt * c->h_sig[i] * c->h_sig_prd[j];
  double V = sg * pt;
  if (!ON(OR_ET_HELIX)) return;
  c->energy[OR_ET_HELIX] += V;
  for (int l = 0; l < 3; ++l) {
    FCA(j - 1)[l] += V * (c->an * p0 * xNO[l] + c->ah * p1 * xHO[l]);
    FCA(j)[l] += V * (c->bn * p0 * xNO[l] + c->bh * p1 * xHO[l]);
    FO(j - 1)[l] += V * (c->cn * p0 * xNO[l] + c->ch * p1 * xHO[l]);
    FO(i)[l] -= V * (p0 * xNO[l] + p1 * xHO[l]);
  }
}

/* helix pair selection, fb.cpp:7843-7868 (j = i + helix_i_diff contiguous
 * on the same chain; the reference also requires the O-N distance below
 * helix_cutoff before calling compute_helix_dtheta_pair) */
static int helix_pair_ok(const or_ctx *c, int i)
{
  int j = i + c->helix_i_diff;
  if (j >= c->n) return 0;
  if (c->chain_no[i] != c->chain_no[j]) return 0;
  if (rsq_of(XO(i), XN(j)) >= c->helix_cutoff_sq) return 0;
  return 1;
}

/* compute_pair, fb.cpp:7677-8135 */
/* ------------------------------------------------------------------ */
/* coeff-toggled terms                                                  */

/* the neighbour list of the list-limited terms: every pair closer than
 * pair_list_cutoff + skin, as a list freshly built by LAMMPS holds
 * (fb.cpp:1805-1813; skin 0 by default: a list rebuilt every step) */
static inline double list_r2(const or_ctx *c)
{
  double r = c->pair_list_cutoff + c->list_skin;
  return r * r;
}

/* get_water_gamma / get_burial_gamma, fb.cpp:2872-2900: a pair with a
 * phosphorylated residue reads the hypercharged table; a phosphorylated
 * residue's burial gammas are those of type 6 */
static inline double water_g(const or_ctx *c, int i, int j, int w, int it, int jt, int k)
{
  if (c->phosph_flag && (c->phosph_map[i] || c->phosph_map[j])) return c->phosph_water_gamma[w][it][jt][k];
  return c->water_gamma[w][it][jt][k];
}
static inline double burial_g(const or_ctx *c, int i, int it, int b)
{
  if (c->phosph_flag && c->phosph_map[i]) return c->burial_gamma[6][b];
  return c->burial_gamma[it][b];
}

/* compute_membrane_potential, fb.cpp:6107-6235 (the cylindrical pore of
 * memb_pore_type 0; the other form is commented out in the reference) */
static void memb_term(or_ctx *c, int i)
{
  const double *xi = c->se[i] == 'G' ? XCA(i) : XCB(i);
  double *fi = c->se[i] == 'G' ? FCA(i) : FCB(i);
  double dx = xi[0] - c->memb_xo[0], dy = xi[1] - c->memb_xo[1], dz = xi[2] - c->memb_xo[2];
  double kb = c->k_bin, a = c->rho0_distor * c->rho0_max, b = c->memb_len / 2;
  double rho = sqrt(dx * dx + dy * dy);
  double rho0 = (c->rho0_max - a) + (a / c->memb_len) * (dz + b);
  double s_per = 0.5 * (1 + tanh(kb * (dz - b)));
  double s_mem = 0.5 * (tanh(kb * (dz + b)) + tanh(kb * (b - dz)));
  double s_cyt = 0.5 * (1 + tanh(kb * (-b - dz)));
  double s_por = 0.5 * (1 - tanh(kb * (rho - rho0)));
  double s_nopor = 1 - s_por;
  int z = c->z_res[i] - 1;
  const double *g = c->g_memb[z];
  /* classes 1 and 3 share the signs, class 2 has them reversed */
  double sg = z == 1 ? -1.0 : 1.0;
  double V = sg * (-g[0] * s_per - g[1] * s_cyt + g[2] * s_mem * s_nopor - g[3] * s_mem * s_por);
  c->energy[OR_ET_MEMB] += c->epsilon * c->k_overall_memb * V;
  double dz_per = 0.5 * kb * (1 - pow(tanh(kb * (dz - b)), 2));
  double dz_mem = -0.5 * kb * pow(tanh(kb * (dz + b)), 2) + 0.5 * kb * pow(tanh(kb * (b - dz)), 2);
  double dz_cyt = -0.5 * kb * (1 - pow(tanh(kb * (-b - dz)), 2));
  double dr1_dz = a / c->memb_len;
  double sech2 = 1 - pow(tanh(kb * (rho - rho0)), 2);
  double dz_s_por = 0.5 * kb * sech2 * dr1_dz;
  double dx_s_por = ((-0.5 * kb * dx) / rho) * sech2, dy_s_por = ((-0.5 * kb * dy) / rho) * sech2;
  double dz_s_por_smem = s_mem * dz_s_por + dz_mem * s_por;
  double dz_s_nopor_smem = s_mem * (-dz_s_por) + dz_mem * s_nopor;
  double dVx = sg * (g[2] * s_mem * (-dx_s_por) - g[3] * s_mem * dx_s_por);
  double dVy = sg * (g[2] * s_mem * (-dy_s_por) - g[3] * s_mem * dy_s_por);
  double dVz = sg * (-g[0] * dz_per - g[1] * dz_cyt + g[2] * dz_s_nopor_smem - g[3] * dz_s_por_smem);
  double k = -c->epsilon * c->k_overall_memb;
  fi[0] += k * dVx; fi[1] += k * dVy; fi[2] += k * dVz;
}

/* compute_vector_fragment_memory_potential, fb.cpp:3994-4077: for the
 * memories covering i and j from i + minSep to the memory's end (at most
 * i + maxSep), no glycine in the target or the memory at i or j: the angle
 * gc between the CA->CB vectors against the memory's (Fragment_Memory::VMf),
 * V = -eps k_vfm exp(-(gc - gf)^2 / (2 sigma^2)) */
static void vfm_term(or_ctx *c, int i)
{
  int n = c->n, maxsep = c->fmg.i_sep[c->fmg.nseq_cl], minsep = c->fmg.i_sep[0];
  double ekw = c->epsilon * c->k_vec_frag_mem;
  for (int ifm = 0; ifm < c->ilen_fm_map[i]; ++ifm) {
    const or_mem *m = &c->mems[c->frag_mem_map[i][ifm]];
    int js = i + minsep, je = m->pos + m->len - 1;
    if (maxsep != -1 && je > i + maxsep) je = i + maxsep;
    if (je >= n) continue; /* reference aborts: "Missing residues in memory potential" */
    for (int j = js; j <= je; ++j) {
      if (c->se[i] == 'G' || c->se[j] == 'G' || m->se[i - m->pos] == 'G' || m->se[j - m->pos] == 'G') continue;
      double vi[3] = {XCB(i)[0] - XCA(i)[0], XCB(i)[1] - XCA(i)[1], XCB(i)[2] - XCA(i)[2]};
      double vj[3] = {XCB(j)[0] - XCA(j)[0], XCB(j)[1] - XCA(j)[1], XCB(j)[2] - XCA(j)[2]};
      double vmsqi = dot3(vi, vi), vmsqj = dot3(vj, vj);
      double vmi = sqrt(vmsqi), vmj = sqrt(vmsqj), vp = dot3(vi, vj);
      double vpn = vp / (vmi * vmj);
      double gc = acos(vpn), gf = m->vmf[(i - m->pos) * m->len + (j - m->pos)];
      double dg = gc - gf;
      double V = -ekw * exp(-dg * dg / (2 * c->vfm_sigma_sq));
      c->energy[OR_ET_VFRAGMEM] += V;
      double force = -V * dg / (c->vfm_sigma_sq * vmi * vmj * sqrt(1 - vpn * vpn));
      double fi[3], fj[3];
      for (int k = 0; k < 3; ++k) {
        fi[k] = force * (vj[k] - vi[k] * vp / vmsqi);
        fj[k] = force * (vi[k] - vj[k] * vp / vmsqj);
      }
      fadd(FCA(i), -1.0, fi); fadd(FCB(i), 1.0, fi);
      fadd(FCA(j), -1.0, fj); fadd(FCB(j), 1.0, fj);
    }
  }
}

/* compute_contact_restraints_potential, fb.cpp:8137-8177: the first entry
 * for (min, max) of the pair; V = -w exp(-dr^2 / (2 sigma^2)) within
 * |dr| < 4.29 sigma; returns the pair force factor (f_i += factor dx) */
static double cr_force(or_ctx *c, int i, int j, double rsq)
{
  int lo = i < j ? i : j, hi = i < j ? j : i;
  for (int q = c->cr_start[lo]; q < c->cr_start[lo + 1]; ++q) {
    if (c->cr_i2[q] != hi) continue;
    double r = sqrt(rsq), dr = r - c->cr_r0[q];
    if (fabs(dr) < c->cr_dr_cutoff) {
      double V = -c->cr_w[q] * exp(-0.5 * dr * dr * c->cr_sigma_sq_inv);
      c->energy[OR_ET_CONT_REST] += V;
      return V * dr * c->cr_sigma_sq_inv / r;
    }
    return 0.0;
  }
  return 0.0;
}

/* compute_solvent_barrier, fb.cpp:6237-6305: two tanh barriers between the
 * CB sites (CA for glycine), radii shifted by residue type when ssb_rad_cor;
 * theta = (tanh k(r - rmin) + tanh k(rmax - r)) / 2, E = eps (k1 theta1 +
 * k2 theta2), zero beyond both rmax + 10/kappa */
static void ssb_term(or_ctx *c, int i, int j)
{
  const double *xi = site(c, i), *xj = site(c, j);
  double dx[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
  double r = sqrt(dot3(dx, dx));
  double rmin1 = c->ssb_rmin1, rmax1 = c->ssb_rmax1, rmin2 = c->ssb_rmin2, rmax2 = c->ssb_rmax2;
  if (c->ssb_rad_cor) {
    double sh = c->ssb_rshift[restype(c->se[i])] + c->ssb_rshift[restype(c->se[j])];
    rmin1 += sh; rmax1 += sh; rmin2 += sh; rmax2 += sh;
  }
  double kp = c->ssb_kappa;
  if (r > rmax1 + 10 / kp && r > rmax2 + 10 / kp) return;
  double tmin1 = tanh(kp * (r - rmin1)), tmax1 = tanh(kp * (rmax1 - r));
  double tmin2 = tanh(kp * (r - rmin2)), tmax2 = tanh(kp * (rmax2 - r));
  double th1 = 0.5 * (tmin1 + tmax1), th2 = 0.5 * (tmin2 + tmax2);
  double e = c->epsilon;
  c->energy[OR_ET_SSB] += e * c->k_solventb1 * th1;
  c->energy[OR_ET_SSB] += e * c->k_solventb2 * th2;
  double f1 = -e * c->k_solventb1 * kp * th1 * (tmax1 - tmin1) / r;
  double f2 = -e * c->k_solventb2 * kp * th2 * (tmax2 - tmin2) / r;
  fadd(fsite(c, i), f1 + f2, dx);
  fadd(fsite(c, j), -(f1 + f2), dx);
}

/* compute_DebyeHuckel_Interaction, fb.cpp:6307-6363: screened Coulomb
 * between the CB sites of charged residues, k++ / k-- / k+- by the signs */
static void dh_term(or_ctx *c, int i, int j)
{
  if (abs(i - j) < c->debye_huckel_min_sep) return;
  double qi = c->charge[i], qj = c->charge[j];
  if (qi == 0 && qj == 0) return;
  const double *xi = site(c, i), *xj = site(c, j);
  double dx[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
  double r = sqrt(dot3(dx, dx));
  double qq = 0.0;
  if (qi > 0.0 && qj > 0.0) qq = c->k_PlusPlus * qi * qj / r;
  else if (qi < 0.0 && qj < 0.0) qq = c->k_MinusMinus * qi * qj / r;
  else if ((qi < 0.0 && qj > 0.0) || (qi > 0.0 && qj < 0.0)) qq = c->k_PlusMinus * qi * qj / r;
  double E = c->epsilon * qq * exp(-c->k_screening * r / c->screening_length);
  c->energy[OR_ET_DH] += E;
  double ft = (E / r) * (1.0 / r + c->k_screening / c->screening_length);
  fadd(fsite(c, i), ft, dx);
  fadd(fsite(c, j), -ft, dx);
}

/* compute_amh_go_model, fb.cpp:3854-3992: for every CA and non-glycine CB
 * atom i, E_i = sum over atoms j of the full neighbour list (CA or
 * non-glycine CB, |i - j| >= minSep or another chain, not censored) with a
 * native distance below rc and (r - rnative)^2 < 27.6 sigma^2
 * (sigma^2 = |i - j|^0.3) of gamma exp(-(r - rnative)^2 / (2 sigma^2));
 * E = -k/2 sum_i E_i^p / a, forces of atom i's pairs scaled by
 * -k p E_i^(p-1) / (2 a).  The neighbour list is taken as every atom pair
 * closer than pair_list_cutoff (a list rebuilt every step without skin). */
static void amhgo_term(or_ctx *c)
{
  int n = c->n;
  double plc2 = list_r2(c), E = 0.0;
  int minsep = c->amhg.i_sep[0];
  for (int i = 0; i < n; ++i) {
    int it = restype(c->se[i]);
    for (int ka = 0; ka < 2; ++ka) {
      if (ka == 1 && c->se[i] == 'G') continue;
      const double *xa = ka ? XCB(i) : XCA(i);
      double *fa = ka ? FCB(i) : FCA(i);
      double fself[3] = {0.0, 0.0, 0.0}, Ei = 0.0;
      memset(c->go_f, 0, sizeof(double) * 6 * n);
      for (int j = 0; j < n; ++j) {
        if (c->censor && (c->censor[(size_t)i * n + j] || c->censor[(size_t)j * n + i])) continue;
        if (!(abs(i - j) >= minsep || c->chain_no[i] != c->chain_no[j])) continue;
        int jt = restype(c->se[j]);
        for (int kb = 0; kb < 2; ++kb) {
          if (kb == 1 && c->se[j] == 'G') continue;
          if (j == i && kb == ka) continue;
          const double *xb = kb ? XCB(j) : XCA(j);
          double dx[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
          double rsq = dot3(dx, dx);
          if (!(rsq < plc2)) continue;
          double rn = amhgo_rnative(c, i, ka ? FM_CB : FM_CA, j, kb ? FM_CB : FM_CA);
          if (!(rn < c->amh_go_rc)) continue;
          double r = sqrt(rsq), dr = r - rn, drsq = dr * dr;
          double s2 = pow(abs(i - j), 0.3);
          if (!(drsq < 27.6 * s2)) continue;
          double g = gm_get2(&c->amhg, it, jt, i, j);
          double Eij = g * exp(-drsq / (2.0 * s2));
          double force = Eij * dr / (s2 * r);
          for (int k = 0; k < 3; ++k) {
            fself[k] += force * dx[k];
            c->go_f[(2 * j + kb) * 3 + k] -= force * dx[k];
          }
          Ei += Eij;
        }
      }
      double factor = -0.5 * c->k_amh_go * c->amh_go_p * pow(Ei, c->amh_go_p - 1) / c->amh_go_norm;
      for (int k = 0; k < 3; ++k) fa[k] += factor * fself[k];
      for (int j = 0; j < n; ++j)
        for (int kb = 0; kb < 2; ++kb) {
          double *g3 = &c->go_f[(2 * j + kb) * 3];
          if (g3[0] == 0.0 && g3[1] == 0.0 && g3[2] == 0.0) continue;
          double *fb = kb ? FCB(j) : FCA(j);
          for (int k = 0; k < 3; ++k) fb[k] += factor * g3[k];
        }
      E += -0.5 * c->k_amh_go * pow(Ei, c->amh_go_p) / c->amh_go_norm;
    }
  }
  c->energy[OR_ET_AMHGO] += E;
}

static void pair_terms(or_ctx *c)
{
  int n = c->n;
  for (int i = 0; i < n; ++i) {
    c->water_ro[i] = c->helix_ro[i] = c->w_sig[i] = c->w_sig_prd[i] = c->h_sig[i] = c->h_sig_prd[i] = 0.0;
    c->water_xi[i] = c->helix_xi_1[i] = c->helix_xi_2[i] = c->burial_force[i] = 0.0;
  }
  /* DL1: densities, fb.cpp:7729-7786 */
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (!(c->chain_no[i] != c->chain_no[j] || abs(i - j) > 1)) continue;
      double rsq = rsq_of(site(c, i), site(c, j)), th, prd;
      if (c->water_flag && rsq > c->w_rmin_theta_sq[0] && rsq < c->w_rmax_theta_sq[0]) {
        wtheta(c, 0, sqrt(rsq), &th, &prd);
        c->water_ro[i] += th; c->water_ro[j] += th;
      }
      if (c->helix_flag && rsq > c->h_rmin_theta_sq && rsq < c->h_rmax_theta_sq) {
        htheta(c, sqrt(rsq), &th, &prd);
        c->helix_ro[i] += th; c->helix_ro[j] += th;
      }
    }
  /* SL: sigma and burial, fb.cpp:7795-7841 */
  for (int i = 0; i < n; ++i) {
    int it = restype(c->se[i]);
    double th;
    if (c->water_flag) {
      th = tanh(c->water_kappa_sigma * (c->water_ro[i] - c->treshold));
      c->w_sig[i] = 0.5 * (1.0 - th);
      c->w_sig_prd[i] = -c->water_kappa_sigma * c->w_sig[i] * (1.0 + th);
    }
    if (c->helix_flag) {
      th = tanh(c->helix_kappa_sigma * (c->helix_ro[i] - c->helix_treshold));
      c->h_sig[i] = 0.5 * (1.0 - th);
      c->h_sig_prd[i] = -c->helix_kappa_sigma * c->h_sig[i] * (1.0 + th);
    }
    if (c->burial_flag) {
      double t[3][2], rho = c->water_ro[i], kb = c->burial_kappa;
      for (int b = 0; b < 3; ++b) {
        t[b][0] = tanh(kb * (rho - c->burial_ro_min[b]));
        t[b][1] = tanh(kb * (c->burial_ro_max[b] - rho));
      }
      if (ON(OR_ET_BURIAL))
        for (int b = 0; b < 3; ++b) c->energy[OR_ET_BURIAL] += -0.5 * c->k_burial * burial_g(c, i, it, b) * (t[b][0] + t[b][1]);
      double bf = 0.0;
      for (int b = 0; b < 3; ++b) bf += burial_g(c, i, it, b) * (t[b][1] * t[b][1] - t[b][0] * t[b][0]);
      c->burial_force[i] = bf * 0.5 * c->k_burial * kb;
    }
  }
  /* helix (i, i+4), fb.cpp:7843-7868 */
  if (c->helix_flag)
    for (int i = 0; i < n; ++i)
      if (helix_pair_ok(c, i)) helix_term(c, i, i + c->helix_i_diff);
  /* DL2: water xi, fb.cpp:7879-7952 */
  if (c->water_flag) {
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        if (!(c->chain_no[i] != c->chain_no[j] || abs(i - j) >= c->contact_cutoff)) continue;
        double rsq = rsq_of(site(c, i), site(c, j));
        int it = restype(c->se[i]), jt = restype(c->se[j]);
        for (int w = 0; w < c->n_wells; ++w) {
          if (!c->well_flag[w]) continue;
          double g0 = water_g(c, i, j, w, it, jt, 0), g1 = water_g(c, i, j, w, it, jt, 1);
          /* the reference predicate !fabs(g0-g1)<delta parses as (!fabs(g0-g1)) < delta */
          if (((double)(!fabs(g0 - g1)) < DELTA) && rsq > c->w_rmin_theta_sq[w] && rsq < c->w_rmax_theta_sq[w]) {
            double th, prd;
            wtheta(c, w, sqrt(rsq), &th, &prd);
            double tg = (g1 - g0) * th;
            c->water_xi[i] += tg * c->w_sig[j];
            c->water_xi[j] += tg * c->w_sig[i];
          }
        }
      }
    for (int i = 0; i < n; ++i) c->water_xi[i] *= c->w_sig_prd[i];
  }
  /* DL3: main loop, fb.cpp:7958-8132 */
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      int same = c->chain_no[i] == c->chain_no[j];
      int dres = j - i;
      if (!same || dres > 1) {
        const double *si = site(c, i), *sj = site(c, j);
        double dx[3] = {si[0] - sj[0], si[1] - sj[1], si[2] - sj[2]};
        double rsq = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
        double force = 0.0, r = 0.0;
        int br = 0;
        int it = restype(c->se[i]), jt = restype(c->se[j]);
        if (c->water_flag) {
          for (int w = 0; w < c->n_wells; ++w) {
            if (!c->well_flag[w]) continue;
            if ((!same || dres >= c->contact_cutoff) && rsq > c->w_rmin_theta_sq[w] && rsq < c->w_rmax_theta_sq[w]) {
              double g0 = water_g(c, i, j, w, it, jt, 0), g1 = water_g(c, i, j, w, it, jt, 1), sg;
              if (fabs(g0 - g1) < DELTA) sg = 0.5 * (g0 + g1);
              else sg = g0 + (g1 - g0) * c->w_sig[i] * c->w_sig[j];
              if (!br) { r = sqrt(rsq); br = 1; }
              double th, prd;
              wtheta(c, w, r, &th, &prd);
              if (ON(OR_ET_WATER)) { c->energy[OR_ET_WATER] += -sg * th; force += sg * prd; }
            }
          }
          if (rsq > c->w_rmin_theta_sq[0] && rsq < c->w_rmax_theta_sq[0] && ON(OR_ET_WATER)) {
            double th, prd;
            if (!br) { r = sqrt(rsq); br = 1; }
            wtheta(c, 0, r, &th, &prd);
            if (fabs(c->water_xi[i]) > DELTA_WATER_XI) force += prd * c->water_xi[i];
            if (fabs(c->water_xi[j]) > DELTA_WATER_XI) force += prd * c->water_xi[j];
          }
        }
        if (c->burial_flag && rsq > c->w_rmin_theta_sq[0] && rsq < c->w_rmax_theta_sq[0] && ON(OR_ET_BURIAL)) {
          double th, prd;
          if (!br) { r = sqrt(rsq); br = 1; }
          wtheta(c, 0, r, &th, &prd);
          force += (c->burial_force[i] + c->burial_force[j]) * prd;
        }
        if (c->helix_flag && rsq > c->h_rmin_theta_sq && rsq < c->h_rmax_theta_sq && ON(OR_ET_HELIX)) {
          int hd = c->helix_i_diff;
          double factor = c->helix_xi_1[i] + c->helix_xi_1[j];
          if (i - hd >= 0) factor += c->helix_xi_2[i - hd];
          if (j - hd >= 0) factor += c->helix_xi_2[j - hd];
          if (fabs(factor) > DELTA_HELIX_XI) {
            double th, prd;
            if (!br) { r = sqrt(rsq); br = 1; }
            htheta(c, r, &th, &prd);
            force += -factor * prd;
          }
        }
        /* contact restraints, fb.cpp:8059-8061 */
        if (c->cont_rest_flag && rsq < c->cr_glob_cutoff_sq && ON(OR_ET_CONT_REST)) force += cr_force(c, i, j, rsq);
        if (force != 0.0) {
          fadd(fsite(c, i), force, dx);
          fadd(fsite(c, j), -force, dx);
        }
      }
      /* DSSP: (O of i, CA of j) and (O of j, CA of i), fb.cpp:8064-8094 */
      if (c->dssp_hdrgn_flag && (!same || dres > 2)) {
        /* i acceptor, j donor */
        if (c->se[j] != 'P' && !is_last(c, i) && !is_first(c, j) && j > 0 &&
            rsq_of(XO(i), XN(j)) < c->dssp_hdrgn_cut_sq)
          dssp_term(c, i, j);
        /* j acceptor, i donor */
        if (c->se[i] != 'P' && !is_last(c, j) && !is_first(c, i) && i > 0 &&
            rsq_of(XO(j), XN(i)) < c->dssp_hdrgn_cut_sq)
          dssp_term(c, j, i);
      }
      /* P_AP: CA-CA pairs, fb.cpp:8096-8103 */
      if (c->p_ap_flag && rsq_of(XCA(i), XCA(j)) < c->pap_cutoff_sq) pap_term(c, i, j);
      /* solvent barrier and Debye-Hueckel on the CA-CA pairs of the list
       * (fb.cpp:8113-8117; the list: CA-CA closer than pair_list_cutoff).
       * A same-chain barrier pair counts with the lower residue first, as
       * the reference's check at fb.cpp:6239 and its list-free loop
       * (fb.cpp:7481) have it. */
      if ((c->ssb_flag || c->huckel_flag) && rsq_of(XCA(i), XCA(j)) < list_r2(c)) {
        if (c->ssb_flag && (!same || dres >= c->ssb_ij_sep) && ON(OR_ET_SSB)) ssb_term(c, i, j);
        if (c->huckel_flag && dres > 1 && ON(OR_ET_DH)) dh_term(c, i, j);
      }
    }
}

/* ------------------------------------------------------------------ */
/* PairExcludedVolume::compute (pev.cpp:50-138) and harmonic bonds      */

static inline int atom_type(const or_ctx *c, int res, int kind)
{
  if (kind == 0) return 1;
  if (kind == 2) return 3;
  return c->se[res] == 'G' ? 5 : 4;
}

/* 1-2 bonded pairs removed by special_bonds fene (deck folding.in:17):
 * CA(i)-CA(i+1), CA(i)-O(i), O(i)-CA(i+1), CA(i)-CB(i) on one chain. */
static int bonded12(const or_ctx *c, int ra, int ka, int rb, int kb)
{
  if (ra > rb || (ra == rb && ka > kb)) { int t = ra; ra = rb; rb = t; t = ka; ka = kb; kb = t; }
  if (ra == rb) return (ka == 0 && (kb == 1 || kb == 2));
  if (rb != ra + 1 || c->chain_no[ra] != c->chain_no[rb]) return 0;
  return (ka == 0 && kb == 0) || (ka == 2 && kb == 0);
}

static void ev_terms(or_ctx *c, long long *list, int cap, int *count)
{
  int n = c->n, p = c->ev_p;
  double sign = (p % 2 == 0) ? 1.0 : -1.0;
  const double *X[3] = {c->xca, c->xcb, c->xo};
  double *F[3] = {c->fca, c->fcb, c->fo};
  int na = 3 * n;
  for (int a = 0; a < na; ++a) {
    int ra = a / 3, ka = a % 3;
    for (int b = a + 1; b < na; ++b) {
      int rb = b / 3, kb = b % 3;
      if (bonded12(c, ra, ka, rb, kb)) continue;
      int ta = atom_type(c, ra, ka), tb = atom_type(c, rb, kb);
      if (c->ev_lambda[ta][tb] == 0.0) continue; /* contributes exactly zero */
      const double *xa = X[ka] + 3 * ra, *xb = X[kb] + 3 * rb;
      double d[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
      double rsq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double r = sqrt(rsq);
      double rc = (abs(ra - rb) < 5 && c->chain_no[ra] == c->chain_no[rb]) ? c->ev_cut_short[ta][tb] : c->ev_cut_long[ta][tb];
      if (r < rc) {
        if (list) {
          if (*count < cap) list[*count] = (long long)a * na + b;
          (*count)++;
          continue;
        }
        double dr = r - rc;
        double pref = sign * c->ev_lambda[ta][tb] * pow(dr, p - 1);
        double fpair = r > 0.0 ? -pref * p / r : 0.0;
        fadd(F[ka] + 3 * ra, fpair, d);
        fadd(F[kb] + 3 * rb, -fpair, d);
        c->energy[OR_ET_VEXCLUDED] += pref * dr;
      }
    }
  }
}

static void bond_one(or_ctx *c, double *fa, double *fb, const double *xa, const double *xb, double r0)
{
  double d[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
  double r = sqrt(dot3(d, d)), dr = r - r0;
  c->energy[OR_ET_BOND] += c->bond_k * dr * dr;
  double fb_ = r > 0.0 ? -2.0 * c->bond_k * dr / r : 0.0;
  fadd(fa, fb_, d);
  fadd(fb, -fb_, d);
}

static void bond_terms(or_ctx *c)
{
  for (int i = 0; i < c->n; ++i) {
    bond_one(c, FCA(i), FO(i), XCA(i), XO(i), c->bond_r0[1]);
    bond_one(c, FCA(i), FCB(i), XCA(i), XCB(i), c->se[i] == 'G' ? c->bond_r0[4] : c->bond_r0[3]);
    if (!is_last(c, i)) {
      bond_one(c, FCA(i), FCA(i + 1), XCA(i), XCA(i + 1), c->bond_r0[0]);
      bond_one(c, FO(i), FCA(i + 1), XO(i), XCA(i + 1), c->bond_r0[2]);
    }
  }
}

/* ------------------------------------------------------------------ */
/* compute_backbone, fb.cpp:7047-7673 (non-DEBUGFORCES path)            */

int or_compute(or_ctx *c, const double *xca, const double *xcb, const double *xo,
               double *fca, double *fcb, double *fo, double *energy, unsigned mask)
{
  c->xca = xca; c->xcb = xcb; c->xo = xo;
  c->fca = fca; c->fcb = fcb; c->fo = fo;
  c->energy = energy;
  c->mask = mask;
  for (int k = 0; k < OR_N_ET; ++k) energy[k] = 0.0;
  build_virtual(c);
  for (int i = 0; i < c->n; ++i) {
    int first = is_first(c, i), last = is_last(c, i);
    if (c->chain_flag && ON(OR_ET_CHAIN)) chain_term(c, i);
    if (!first && !last && c->chi_flag && c->se[i] != 'G' && ON(OR_ET_CHI)) chi_term(c, i);
    if (c->shake_flag && ON(OR_ET_SHAKE)) shake_term(c, i);
    if (!first && !last && c->rama_flag && c->se[i] != 'G' && ON(OR_ET_RAMA)) rama_term(c, i);
    if (c->memb_flag && ON(OR_ET_MEMB)) memb_term(c, i);
    if (c->frag_mem_flag && ON(OR_ET_FRAGMEM)) fm_term(c, i);
    if (c->vec_frag_mem_flag && ON(OR_ET_VFRAGMEM)) vfm_term(c, i);
    if (c->frag_mem_tb_flag && ON(OR_ET_FRAGMEM) && fm_table_terms(c, i)) return -1;
  }
  pair_terms(c);
  if (c->amh_go_flag && ON(OR_ET_AMHGO)) amhgo_term(c);
  if (ON(OR_ET_VEXCLUDED)) ev_terms(c, NULL, 0, NULL);
  if (ON(OR_ET_BOND)) bond_terms(c);
  for (int k = 1; k < OR_N_ET; ++k) energy[OR_ET_TOTAL] += energy[k];
  return 0;
}

/* ------------------------------------------------------------------ */
/* integer lists                                                        */

static int cmp_ll(const void *a, const void *b)
{
  long long x = *(const long long *)a, y = *(const long long *)b;
  return x < y ? -1 : (x > y);
}

int or_list(or_ctx *c, int kind, int arg, const double *xca, const double *xcb, const double *xo,
            long long *out, int cap)
{
  c->xca = xca; c->xcb = xcb; c->xo = xo;
  build_virtual(c);
  int n = c->n, cnt = 0;
#define PUSH(v) do { if (cnt < cap) out[cnt] = (v); cnt++; } while (0)
  if (kind == 5) {
    ev_terms(c, out, cap, &cnt);
  } else if (kind == 4) {
    if (c->helix_flag)
      for (int i = 0; i < n; ++i)
        if (helix_pair_ok(c, i) && rNO(c, i, i + c->helix_i_diff) <= c->helix_cutoff) PUSH((long long)i * n + i + c->helix_i_diff);
  } else {
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        int same = c->chain_no[i] == c->chain_no[j], dres = j - i;
        double rsq = rsq_of(site(c, i), site(c, j));
        if (kind == 0 && c->water_flag && (!same || dres > 1) && arg < c->n_wells &&
            rsq > c->w_rmin_theta_sq[arg] && rsq < c->w_rmax_theta_sq[arg]) PUSH((long long)i * n + j);
        if (kind == 1 && c->helix_flag && (!same || dres > 1) && rsq > c->h_rmin_theta_sq && rsq < c->h_rmax_theta_sq)
          PUSH((long long)i * n + j);
        if (kind == 6 && c->water_flag && (!same || dres >= c->contact_cutoff) && arg < c->n_wells && c->well_flag[arg]) {
          int it = restype(c->se[i]), jt = restype(c->se[j]);
          double g0 = water_g(c, i, j, arg, it, jt, 0), g1 = water_g(c, i, j, arg, it, jt, 1);
          if (((double)(!fabs(g0 - g1)) < DELTA) && rsq > c->w_rmin_theta_sq[arg] && rsq < c->w_rmax_theta_sq[arg])
            PUSH((long long)i * n + j);
        }
        if (kind == 2 && c->dssp_hdrgn_flag && (!same || dres > 2)) {
          if (c->se[j] != 'P' && !is_last(c, i) && !is_first(c, j) && rsq_of(XO(i), XN(j)) < c->dssp_hdrgn_cut_sq &&
              !(rNO(c, i, j) > c->dssp_hdrgn_cut)) PUSH((long long)i * n + j);
          if (c->se[i] != 'P' && !is_last(c, j) && !is_first(c, i) && rsq_of(XO(j), XN(i)) < c->dssp_hdrgn_cut_sq &&
              !(rNO(c, j, i) > c->dssp_hdrgn_cut)) PUSH((long long)j * n + i);
        }
        if (kind == 3 && c->p_ap_flag && rsq_of(XCA(i), XCA(j)) < c->pap_cutoff_sq) PUSH((long long)i * n + j);
      }
  }
#undef PUSH
  if (cnt > cap) return -cnt;
  qsort(out, (size_t)cnt, sizeof(long long), cmp_ll);
  return cnt;
}

/* ------------------------------------------------------------------ */
/* fragment-memory introspection                                        */

int or_branch_counts(or_ctx *c, long long *out, int n, int reset)
{
  int m = n < OR_N_BRANCH ? n : OR_N_BRANCH;
  for (int k = 0; k < m; ++k) out[k] = c->branch[k];
  if (reset) memset(c->branch, 0, sizeof c->branch);
  return OR_N_BRANCH;
}

int or_fm_count(const or_ctx *c) { return c->n_mems; }

const char *or_last_error(const or_ctx *c) { return c->last_err; }

/* table mode: grid size and neighbour count (0 without table mode) */
void or_fm_table_dims(const or_ctx *c, int *tb_size, int *tb_nbrs)
{
  *tb_size = c->frag_mem_tb_flag ? c->tb_size : 0;
  *tb_nbrs = c->frag_mem_tb_flag ? c->tb_nbrs : 0;
}

/* curve itb (energy, force on the grid): 1 when it exists, else 0 */
int or_fm_table(const or_ctx *c, int itb, double *e, double *f)
{
  if (!c->frag_mem_tb_flag || itb < 0 || (size_t)itb >= (size_t)4 * c->n * c->tb_nbrs || !c->tb_e[itb]) return 0;
  memcpy(e, c->tb_e[itb], (size_t)c->tb_size * sizeof(double));
  memcpy(f, c->tb_f[itb], (size_t)c->tb_size * sizeof(double));
  return 1;
}
int or_fm_info(const or_ctx *c, int k, int *pos, int *len, double *weight)
{
  if (k < 0 || k >= c->n_mems) return -1;
  *pos = c->mems[k].pos; *len = c->mems[k].len; *weight = c->mems[k].weight;
  return 0;
}
double or_fm_rf(const or_ctx *c, int k, int ires, int iatom, int jres, int jatom)
{
  return mem_rf(&c->mems[k], ires, iatom, jres, jatom);
}
char or_fm_se(const or_ctx *c, int k, int resno) { return c->mems[k].se[resno - c->mems[k].pos]; }
double or_fm_gamma(const or_ctx *c, int it, int jt, int ires, int jres) { return gm_get2(&c->fmg, it, jt, ires, jres); }
int or_fm_minsep(const or_ctx *c) { return c->fmg.i_sep[0]; }
int or_fm_maxsep(const or_ctx *c) { return c->fmg.i_sep[c->fmg.nseq_cl]; }
int or_fm_map(const or_ctx *c, int i, int *out, int cap)
{
  if (!c->ilen_fm_map) return 0;
  int n = c->ilen_fm_map[i];
  for (int k = 0; k < n && k < cap; ++k) out[k] = c->frag_mem_map[i][k];
  return n;
}

/* ------------------------------------------------------------------ */
/* Langevin dynamics: LAMMPS fix nve + fix langevin, real units.        */
/* Counter hash: splitmix64 finaliser over (seed, replica, step, atom,  */
/* component); uniform in [0,1) from the top 53 bits.                   */

static inline uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static inline double urand(uint64_t seed, int replica, long long step, int atom, int comp)
{
  uint64_t h = mix64(seed ^ 0x9e3779b97f4a7c15ULL);
  h = mix64(h ^ ((uint64_t)(uint32_t)replica << 32 | (uint32_t)comp));
  h = mix64(h ^ (uint64_t)step);
  h = mix64(h ^ (uint64_t)(uint32_t)atom);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

#define FTM2V (1.0 / 48.88821291 / 48.88821291)
#define MVV2E (48.88821291 * 48.88821291)
#define BOLTZ 0.0019872067

static inline double atom_mass(const or_ctx *c, int res, int kind)
{
  if (kind == 0) return 27.0;
  if (kind == 2) return 28.0;
  return c->se[res] == 'G' ? 2.0 : 60.0;
}

int or_langevin_run(or_ctx *c, int nsteps, double dt, double t_start, double t_stop, double damp,
                    unsigned long long seed, int replica, long long step0, double *x, double *v, double *f,
                    double *energy)
{
  int n = c->n;
  int na = 3 * n;
  double dtf = 0.5 * dt * FTM2V;
  unsigned all = 0xffffffffu;
  for (int s = 0; s < nsteps; ++s) {
    long long step = step0 + s;
    /* initial_integrate (fix nve) */
    for (int a = 0; a < na; ++a) {
      int kind = a / n, res = a % n;
      double m = atom_mass(c, res, kind);
      double dtfm = dtf / m;
      for (int l = 0; l < 3; ++l) {
        v[3 * a + l] += dtfm * f[3 * a + l];
        x[3 * a + l] += dt * v[3 * a + l];
      }
    }
    /* force */
    for (int k = 0; k < 3 * na; ++k) f[k] = 0.0;
    if (or_compute(c, x, x + 3 * n, x + 6 * n, f, f + 3 * n, f + 6 * n, energy, all)) return -1;
    /* fix langevin post_force: target T ramps linearly over the run */
    double delta = nsteps > 0 ? (double)(s + 1) / (double)nsteps : 0.0;
    double T = t_start + delta * (t_stop - t_start);
    double tsqrt = sqrt(T);
    for (int a = 0; a < na; ++a) {
      int kind = a / n, res = a % n;
      double m = atom_mass(c, res, kind);
      double g1 = -m / damp / FTM2V;
      double g2 = sqrt(m) * sqrt(24.0 * BOLTZ / damp / dt / MVV2E) / FTM2V * tsqrt;
      for (int l = 0; l < 3; ++l) f[3 * a + l] += g1 * v[3 * a + l] + g2 * (urand(seed, replica, step, a, l) - 0.5);
    }
    /* final_integrate */
    for (int a = 0; a < na; ++a) {
      int kind = a / n, res = a % n;
      double dtfm = dtf / atom_mass(c, res, kind);
      for (int l = 0; l < 3; ++l) v[3 * a + l] += dtfm * f[3 * a + l];
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* order parameters: compute qwolynes and compute qonuchic              */

/* ComputeQWolynes::compute_scalar (src/compute_q_wolynes.cpp:171-236):
 * CA pairs i < j with |i - j| > sep, sigma = (1 + |i - j|)^sexp,
 * q = sum exp(-(r - rn_ij)^2 / (2 sigma^2)) * 2 / ((n - sep)(n - sep - 1)),
 * rn row-major n x n (readNativeDistances, 152-167).  The reference's
 * periodic unwrap is overwritten by the stored coordinate (198-214): the
 * coordinates are used as given. */
double or_q_wolynes(int n, const double *xca, const double *rn, int sep, double sexp)
{
  double q = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (!(abs(j - i) > sep)) continue;
      double s = pow(1 + abs(i - j), sexp);
      double dx = xca[3 * i] - xca[3 * j], dy = xca[3 * i + 1] - xca[3 * j + 1], dz = xca[3 * i + 2] - xca[3 * j + 2];
      double r = sqrt(pow(dx, 2) + pow(dy, 2) + pow(dz, 2));
      q += exp(-pow(r - rn[(size_t)i * n + j], 2) / (2.0 * pow(s, 2)));
    }
  return q * (2.0 / (((double)n - (double)sep) * ((double)n - ((double)sep + 1.0))));
}

/* ComputeQOnuchic (src/compute_q_onuchic.cpp:150-317).  type 0 cutoff,
 * 1 cutoff/gauss: native pairs are j - i > 3 with |x_i - x_j|^2 < rc^2 in
 * the native coordinates (nat, n x 3); type 2 shadow, 3 shadow/gauss: the
 * contact map's (i, j, rn) entries (0-based; a later entry of a pair
 * replaces an earlier one, every entry counts in the normalisation).
 * Cutoff types count pairs with r^2 < (factor rn)^2; the gauss types add
 * exp(-(r - rn)^2 / (2 (1 + |i - j|)^0.3)).  Pairs are visited i < j (the
 * atom order of a data file), so a map entry with i > j counts only in the
 * normalisation.  q /= (number of native pairs or map entries). */
double or_q_onuchic(int n, const double *xca, int type, const double *nat, double rc, int nmap, const int *mi,
                    const int *mj, const double *mrn, double factor)
{
  unsigned char *isn = (unsigned char *)calloc((size_t)n * n, 1);
  double *rsq = (double *)calloc((size_t)n * n, sizeof(double));
  double norm = 0.0;
  if (type == 0 || type == 1) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        double dx = nat[3 * j] - nat[3 * i], dy = nat[3 * j + 1] - nat[3 * i + 1], dz = nat[3 * j + 2] - nat[3 * i + 2];
        size_t k = (size_t)i * n + j;
        rsq[k] = dx * dx + dy * dy + dz * dz;
        if (rsq[k] < rc * rc && j - i > 3) { isn[k] = 1; norm += 1.0; }
        if (type == 0 && isn[k]) rsq[k] *= factor * factor;
      }
  } else {
    for (int e = 0; e < nmap; ++e) {
      size_t k = (size_t)mi[e] * n + mj[e];
      isn[k] = 1;
      rsq[k] = mrn[e] * mrn[e];
      if (type == 2) rsq[k] *= factor * factor;
      norm += 1.0;
    }
  }
  double q = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      size_t k = (size_t)i * n + j;
      if (!(isn[k] && abs(j - i) > 3)) continue;
      double dx = xca[3 * i] - xca[3 * j], dy = xca[3 * i + 1] - xca[3 * j + 1], dz = xca[3 * i + 2] - xca[3 * j + 2];
      double r2 = dx * dx + dy * dy + dz * dz;
      if (type == 0 || type == 2) {
        if (r2 < rsq[k]) q += 1.0;
      } else {
        double s2 = pow(1 + abs(i - j), 2.0 * 0.15);
        q += exp(-pow(sqrt(r2) - sqrt(rsq[k]), 2) / (2.0 * s2));
      }
    }
  free(isn);
  free(rsq);
  return q * (1 / norm);
}

/* ------------------------------------------------------------------ */
/* the tertiary frustratometer (fb.cpp:5057-5412, 5444-5600)           */

/* compute_water_energy (fb.cpp:5444-5476): wells 0 (direct, the mean of
 * its two gammas) and 1 (mediated, protein/water gammas by the two
 * residues' sigma) at distance r, densities rho_i, rho_j */
static double tf_water(const or_ctx *c, double r, int i, int j, int it, int jt, double ri, double rj)
{
  double g0d = water_g(c, i, j, 0, it, jt, 0), g1d = water_g(c, i, j, 0, it, jt, 1);
  double gpm = water_g(c, i, j, 1, it, jt, 0), gwm = water_g(c, i, j, 1, it, jt, 1);
  double sw = 0.25 * (1.0 - tanh(c->water_kappa_sigma * (ri - c->treshold))) * (1.0 - tanh(c->water_kappa_sigma * (rj - c->treshold)));
  double sp = 1.0 - sw;
  double sgd = (g0d + g1d) / 2, sgm = sp * gpm + sw * gwm;
  double k = c->water_kappa;
  double td = 0.25 * (1.0 + tanh(k * (r - c->well_r_min[0]))) * (1.0 + tanh(k * (c->well_r_max[0] - r)));
  double tm = 0.25 * (1.0 + tanh(k * (r - c->well_r_min[1]))) * (1.0 + tanh(k * (c->well_r_max[1] - r)));
  return -(sgd * td + sgm * tm);
}

/* compute_burial_energy (fb.cpp:5478-5500) */
static double tf_burial(const or_ctx *c, int i, int it, double rho)
{
  double e = 0.0, kb = c->burial_kappa;
  for (int b = 0; b < 3; ++b)
    e += -0.5 * c->k_burial * burial_g(c, i, it, b) * (tanh(kb * (rho - c->burial_ro_min[b])) + tanh(kb * (c->burial_ro_max[b] - rho)));
  return e;
}

/* compute_electrostatic_energy (fb.cpp:5502-5548): charges by residue type
 * (R, K +1; D, E -1), nothing below the minimum separation of i, j */
static double tf_elec(const or_ctx *c, double r, int i, int j, int it, int jt)
{
  static const char one[20] = {'A', 'R', 'N', 'D', 'C', 'Q', 'E', 'G', 'H', 'I', 'L', 'K', 'M', 'F', 'P', 'S', 'T', 'W', 'Y', 'V'};
  if (abs(i - j) < c->debye_huckel_min_sep) return 0.0;
  double qi, qj;
  if (one[it] == 'R' || one[it] == 'K') qi = 1.0;
  else if (one[it] == 'D' || one[it] == 'E') qi = -1.0;
  else return 0.0;
  if (one[jt] == 'R' || one[jt] == 'K') qj = 1.0;
  else if (one[jt] == 'D' || one[jt] == 'E') qj = -1.0;
  else return 0.0;
  double k = (qi > 0 && qj > 0) ? c->k_PlusPlus : ((qi < 0 && qj < 0) ? c->k_MinusMinus : c->k_PlusMinus);
  return c->epsilon * (k * qi * qj / r) * exp(-c->k_screening * r / c->screening_length);
}

static double tf_dist(const or_ctx *c, int i, int j)
{
  return sqrt(rsq_of(site(c, i), site(c, j)));
}

/* mean and standard deviation (population) of the decoy energies
 * (fb.cpp:5413-5442) */
static void tf_stats(const double *e, int m, double *st)
{
  double mean = 0.0, var = 0.0;
  for (int k = 0; k < m; ++k) mean += e[k];
  mean /= (double)m;
  for (int k = 0; k < m; ++k) var += (e[k] - mean) * (e[k] - mean);
  st[0] = mean;
  st[1] = sqrt(var / (double)m);
}

/* the mutational environment of residue a (its contacts k other than x,
 * y) with a's type at and residue index ar (fb.cpp:5211-5239) */
static double tf_env(const or_ctx *c, int a, int ar, int at, double rho_a, int x, int y, const double *rho)
{
  double e = 0.0;
  for (int k = 0; k < c->n; ++k) {
    if (k == x || k == y) continue;
    double rak = tf_dist(c, a, k);
    int kt = restype(c->se[k]);
    if (rak < c->tf_cutoff) e += tf_water(c, rak, ar, k, at, kt, rho_a, rho[k]);
    if (c->huckel_flag) e += tf_elec(c, rak, ar, k, at, kt);
  }
  return e;
}

/* one call of the frustratometer at the coordinates set by or_compute
 * (densities from the well-0 pairs as cWell::compute_ro, sml.h:630-650).
 * Rows of 11 doubles: i, j, i_chain, j_chain, r_ij, rho_i, rho_j, native,
 * decoy mean, decoy std, frustration index (1-based residue and chain
 * numbers; j = 0 and j_chain = 0 in singleresidue mode).  Returns the rows
 * written, or -needed. */
/* the densities the frustratometers read (get_residue_density,
 * fb.cpp:5579-5583: the well's ro, the well-0 sum of cWell::compute_ro,
 * sml.h:630-650) */
static double *tf_rho(or_ctx *c)
{
  int n = c->n;
  double *rho = (double *)calloc((size_t)n, sizeof(double));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      if (!(c->chain_no[i] != c->chain_no[j] || abs(j - i) > 1)) continue;
      double r2 = rsq_of(site(c, i), site(c, j));
      if (r2 < c->w_rmin_theta_sq[0] || r2 > c->w_rmax_theta_sq[0]) continue;
      double th, prd;
      wtheta(c, 0, sqrt(r2), &th, &prd);
      rho[i] += th;
    }
  return rho;
}

int or_tert_frust(or_ctx *c, const double *xca, const double *xcb, const double *xo, double *rows, int cap)
{
  int n = c->n;
  c->xca = xca; c->xcb = xcb; c->xo = xo;
  double *rho = tf_rho(c);
  uint32_t rng0[34];
  int rngk0 = c->rng_k, done0 = c->tf_done;
  double st0[2] = {c->tf_stats[0], c->tf_stats[1]};
  memcpy(rng0, c->rng, sizeof rng0);
  int m = c->tf_ndecoys, nr = 0;
  double *dec = (double *)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
  double cut = c->tf_cutoff;
#define TF_ROW(a0, a1, a2, a3, a4, a5, a6, a7, a8, a9, a10)                                       \
  do {                                                                                        \
    if (nr < cap) {                                                                           \
      double *q = rows + 11 * (size_t)nr;                                                     \
      q[0] = a0; q[1] = a1; q[2] = a2; q[3] = a3; q[4] = a4; q[5] = a5; q[6] = a6; q[7] = a7; \
      q[8] = a8; q[9] = a9; q[10] = a10;                                                      \
    }                                                                                         \
    nr++;                                                                                     \
  } while (0)
  if (c->tf_mode == 2) {
    for (int i = 0; i < n; ++i) {
      int it = restype(c->se[i]);
      double e = tf_burial(c, i, it, rho[i]);
      /* compute_singleresidue_native_ixn (fb.cpp:5346-5393) */
      for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        double r = tf_dist(c, i, j);
        int jt = restype(c->se[j]);
        if (r < cut && (abs(i - j) >= c->contact_cutoff || c->chain_no[i] != c->chain_no[j]))
          e += tf_water(c, r, i, j, it, jt, rho[i], rho[j]);
        if (c->huckel_flag) e += tf_elec(c, r, i, j, it, jt);
      }
      for (int d = 0; d < m; ++d) {
        int dt = restype(c->se[or_rng_next(c) % n]);
        double ed = tf_burial(c, i, dt, rho[i]);
        for (int j = 0; j < n; ++j) {
          if (j == i) continue;
          double r = tf_dist(c, i, j);
          int jt = restype(c->se[j]);
          if (r < cut && (abs(i - j) >= c->contact_cutoff || c->chain_no[i] != c->chain_no[j]))
            ed += tf_water(c, r, i, j, dt, jt, rho[i], rho[j]);
          if (c->huckel_flag) ed += tf_elec(c, r, i, j, dt, jt);
        }
        dec[d] = ed;
      }
      double st[2];
      tf_stats(dec, m, st);
      TF_ROW(i + 1, 0, c->chain_no[i], 0, 0.0, rho[i], 0.0, e, st[0], st[1], (st[0] - e) / st[1]);
    }
  } else {
    for (int i = 0; i < n; ++i) {
      int it = restype(c->se[i]);
      for (int j = i + 1; j < n; ++j) {
        int jt = restype(c->se[j]);
        double rij = tf_dist(c, i, j);
        if (!(rij < cut && (abs(i - j) >= c->contact_cutoff || c->chain_no[i] != c->chain_no[j]))) continue;
        /* compute_native_ixn, fb.cpp:5191-5247 */
        double e = tf_water(c, rij, i, j, it, jt, rho[i], rho[j]) + tf_burial(c, i, it, rho[i]) + tf_burial(c, j, jt, rho[j]);
        if (c->huckel_flag) e += tf_elec(c, rij, i, j, it, jt);
        if (c->tf_mode == 1) e += tf_env(c, i, i, it, rho[i], i, j, rho) + tf_env(c, j, j, jt, rho[j], i, j, rho);
        /* compute_decoy_ixns, fb.cpp:5249-5344 */
        if (c->tf_mode == 1 || !c->tf_done) {
          for (int d = 0; d < m; ++d) {
            double r = rij, ri = rho[i], rj = rho[j];
            if (c->tf_mode == 0) {
              int a = or_rng_next(c) % n, b = or_rng_next(c) % n;
              r = tf_dist(c, a, b);
              while (r > cut || a == b) {
                a = or_rng_next(c) % n;
                b = or_rng_next(c) % n;
                r = tf_dist(c, a, b);
              }
              a = or_rng_next(c) % n;
              b = or_rng_next(c) % n;
              ri = rho[a];
              rj = rho[b];
            }
            int a = or_rng_next(c) % n, b = or_rng_next(c) % n;
            int at = restype(c->se[a]), bt = restype(c->se[b]);
            double ed = tf_water(c, r, a, b, at, bt, ri, rj) + tf_burial(c, a, at, ri) + tf_burial(c, b, bt, rj);
            if (c->huckel_flag) ed += tf_elec(c, r, a, b, at, bt);
            if (c->tf_mode == 1) ed += tf_env(c, i, a, at, ri, i, j, rho) + tf_env(c, j, b, bt, rj, i, j, rho);
            dec[d] = ed;
          }
          tf_stats(dec, m, c->tf_stats);
          if (c->tf_mode == 0) c->tf_done = 1;
        }
        TF_ROW(i + 1, j + 1, c->chain_no[i], c->chain_no[j], rij, rho[i], rho[j], e, c->tf_stats[0], c->tf_stats[1],
               (c->tf_stats[0] - e) / c->tf_stats[1]);
      }
    }
  }
#undef TF_ROW
  free(rho);
  free(dec);
  if (nr > cap) { /* a call that did not fit leaves the stream where it was */
    memcpy(c->rng, rng0, sizeof rng0);
    c->rng_k = rngk0;
    c->tf_done = done0;
    c->tf_stats[0] = st0[0];
    c->tf_stats[1] = st0[1];
  }
  return nr <= cap ? nr : -nr;
}

/* ------------------------------------------------------------------ */
/* the n-mer frustratometer (fb.cpp:5600-6061)                          */

/* compute_singleresidue_native_ixn with nmercalc (fb.cpp:5346-5393): the
 * burial of residue i with type it, its water with the residues k < i
 * within the cutoff (and the contact separation, or another chain), and
 * its Debye-Hueckel with every k < i */
static double nf_single(const or_ctx *c, int i, int it, const double *rho)
{
  double w = 0.0, b = tf_burial(c, i, it, rho[i]), el = 0.0;
  for (int k = 0; k < c->n; ++k) {
    if (k == i || k > i) continue;
    double r = tf_dist(c, i, k);
    int kt = restype(c->se[k]);
    if (r < c->nf_cutoff && (abs(i - k) >= c->contact_cutoff || c->chain_no[i] != c->chain_no[k]))
      w += tf_water(c, r, i, k, it, kt, rho[i], rho[k]);
    if (c->huckel_flag) el += tf_elec(c, r, i, k, it, kt);
  }
  return w + b + el;
}

/* the energy of n-mers a (types from residues sa + da*o) and b (types
 * from sb + db*o): their burials, then the water of every (a_o, b_p)
 * pair, a outer (compute_nmer_native_ixn, fb.cpp:5947-5988; the decoys
 * 6028-6056 and the traps 5822-5853 with other type sources) */
static double nf_pair(const or_ctx *c, int a, int sa, int da, int b, int sb, int db, const double *rho)
{
  int s = c->nf_size;
  double e = 0.0;
  for (int o = 0; o < s; ++o) e += tf_burial(c, a + o, restype(c->se[sa + da * o]), rho[a + o]);
  for (int o = 0; o < s; ++o) e += tf_burial(c, b + o, restype(c->se[sb + db * o]), rho[b + o]);
  for (int o = 0; o < s; ++o) {
    int at = restype(c->se[sa + da * o]);
    for (int p = 0; p < s; ++p) {
      int bt = restype(c->se[sb + db * p]);
      e += tf_water(c, tf_dist(c, a + o, b + p), a + o, b + p, at, bt, rho[a + o], rho[b + p]);
    }
  }
  return e;
}

/* secondary-structure letters of an n-mer from the ssweight columns
 * (get_nmer_secondary_structure, fb.cpp:5931-5945) */
static void nf_ss(const or_ctx *c, int st, char *ss)
{
  for (int o = 0; o < c->nf_size; ++o) {
    int e = c->aps[4][st + o] == 1.0, h = c->aps[3][st + o] == 1.0;
    ss[o] = (e && h) ? '!' : (h ? 'H' : (e ? 'E' : '-'));
  }
}

/* one call at the coordinates given.  Rows of 7 doubles: pairwise i, j
 * (1-based starts), contacts, native, decoy mean, decoy std, frustration
 * index; singlenmer i, 0, 0, native, mean, std, index.  With the trap flag
 * (pairwise), trap records of 7 doubles in the reference's order (per row,
 * compute_nmer_traps(i, j) then (j, i), fb.cpp:5620-5623): row, first and
 * second n-mer (1-based), k (1-based), backward, threshold, trap energy.
 * Returns the rows (or -needed); *ntr the trap records (or -needed). */
int or_nmer_frust(or_ctx *c, const double *xca, const double *xcb, const double *xo, double *rows, int cap,
                  double *traps, int tcap, int *ntr)
{
  int n = c->n, s = c->nf_size, m = c->nf_ndecoys, nr = 0, nt = 0;
  c->xca = xca; c->xcb = xcb; c->xo = xo;
  double *rho = tf_rho(c);
  uint32_t rng0[34];
  int rngk0 = c->rng_k;
  memcpy(rng0, c->rng, sizeof rng0);
  double *dec = (double *)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
  char *ssf = (char *)calloc((size_t)s + 1, 1), *ssk = (char *)calloc((size_t)s + 1, 1);
#define NF_ROW(a0, a1, a2, a3, a4, a5, a6)                                                   \
  do {                                                                                     \
    if (nr < cap) {                                                                        \
      double *q = rows + 7 * (size_t)nr;                                                   \
      q[0] = a0; q[1] = a1; q[2] = a2; q[3] = a3; q[4] = a4; q[5] = a5; q[6] = a6;         \
    }                                                                                      \
    nr++;                                                                                  \
  } while (0)
  if (c->nf_mode == 1) {
    /* compute_singlenmer_frust, fb.cpp:5659-5706 */
    for (int i = 0; i < n - s + 1; ++i) {
      double e = 0.0;
      for (int j = i; j < i + s; ++j) e += nf_single(c, j, restype(c->se[j]), rho);
      for (int d = 0; d < m; ++d) {
        int jr = or_rng_next(c) % n;
        while (jr + s > n) jr = or_rng_next(c) % n;
        double ed = 0.0;
        for (int j = i; j < i + s; ++j) ed += nf_single(c, j, restype(c->se[jr + j - i]), rho);
        dec[d] = ed;
      }
      double st[2];
      tf_stats(dec, m, st);
      NF_ROW(i + 1, 0, 0, e, st[0], st[1], (st[0] - e) / st[1]);
    }
  } else {
    /* compute_nmer_frust, fb.cpp:5600-5657 */
    for (int i = 0; i < n - s; ++i)
      for (int j = i + 1; j < n - s; ++j) {
        int cnt = 0; /* compute_nmer_contacts, fb.cpp:5895-5919 */
        for (int a = i; a < i + s; ++a)
          for (int b = j; b < j + s; ++b)
            if (abs(a - b) >= c->contact_cutoff && tf_dist(c, a, b) < c->nf_cutoff) cnt++;
        if (!(cnt >= c->nf_contacts && j - i >= s)) continue;
        double e = nf_pair(c, i, i, 1, j, j, 1, rho);
        for (int d = 0; d < m; ++d) { /* compute_nmer_decoy_ixns, fb.cpp:5990-6060 */
          int ir = or_rng_next(c) % n, jr = or_rng_next(c) % n;
          if (ir > jr) { int t = ir; ir = jr; jr = t; }
          while (ir + s > n || jr + s > n || jr - ir < s) {
            ir = or_rng_next(c) % n;
            jr = or_rng_next(c) % n;
            if (ir > jr) { int t = ir; ir = jr; jr = t; }
          }
          dec[d] = nf_pair(c, i, ir, 1, j, jr, 1, rho);
        }
        double st[2];
        tf_stats(dec, m, st);
        if (c->nf_trap) {
          /* compute_nmer_traps (fb.cpp:5786-5893) for (i, j) and (j, i):
             the second n-mer's types come from residues o (forward) or
             s - o (backward) -- not from k, so the energy is k's for every
             k that passes the overlap and secondary-structure tests */
          double thr = e - c->nf_trap_sigma * st[1];
          for (int dir = 0; dir < 2; ++dir) {
            int f = dir ? j : i, g = dir ? i : j;
            nf_ss(c, g, ssf);
            for (int bw = 0; bw < 2; ++bw) {
              double et = nf_pair(c, f, f, 1, g, bw ? s : 0, bw ? -1 : 1, rho);
              for (int k = 0; k < n - s; ++k) {
                if (abs(k - g) <= s || (abs(k - f) <= s && f != k)) continue;
                nf_ss(c, k, ssk);
                int sd = 0;
                for (int o = 0; o < s; ++o) sd += ssf[o] != ssk[o];
                if (sd > s * (1 - c->nf_ss_frac)) continue;
                if (et < thr) {
                  if (nt < tcap) {
                    double *q = traps + 7 * (size_t)nt;
                    q[0] = nr; q[1] = f + 1; q[2] = g + 1; q[3] = k + 1; q[4] = bw; q[5] = thr; q[6] = et;
                  }
                  nt++;
                }
              }
            }
          }
        }
        NF_ROW(i + 1, j + 1, cnt, e, st[0], st[1], (st[0] - e) / st[1]);
      }
  }
#undef NF_ROW
  free(rho);
  free(dec);
  free(ssf);
  free(ssk);
  if (nr > cap || nt > tcap) { /* a call that did not fit leaves the stream where it was */
    memcpy(c->rng, rng0, sizeof rng0);
    c->rng_k = rngk0;
  }
  *ntr = nt <= tcap ? nt : -nt;
  return nr <= cap ? nr : -nr;
}

/* ------------------------------------------------------------------ */
/* the fragment frustratometer (fb.cpp:4431-4832, 7537-7576)            */

/* compute_decoy_memory_potential (fb.cpp:4431-4563): residue i's memory
 * energies with the library mems (map, ilen) at the current coordinates,
 * added to E[i] and E[j]; with dec != NULL (compute_generated_decoy_
 * energies, 4691-4832) the distances are the decoy structure's instead
 * (0 where the decoy does not cover the residue: its Rf returns 0 there
 * and the reference does not check) */
static int ff_energy(const or_ctx *c, const or_mem *mems, int **map, const int *ilen, const or_mem *dec, double *E)
{
  int n = c->n, minsep = c->fmg.i_sep[0], maxsep = c->fmg.i_sep[c->fmg.nseq_cl];
  const int ity[4] = {FM_CA, FM_CA, FM_CB, FM_CB}, jty[4] = {FM_CA, FM_CB, FM_CA, FM_CB};
  for (int i = 0; i < n; ++i) {
    int itype = restype(c->se[i]);
    for (int ifm = 0; ifm < ilen[i]; ++ifm) {
      const or_mem *m = &mems[map[i][ifm]];
      int js = i + minsep, je = m->pos + m->len - 1;
      if (maxsep != -1 && je > i + maxsep) je = i + maxsep;
      double ekw = c->epsilon * c->k_frag_mem * m->weight;
      if (je >= n) return -1; /* "Missing residues in decoy memory potential" */
      for (int j = js; j <= je; ++j) {
        int jtype = restype(c->se[j]);
        if (c->chain_no[i] != c->chain_no[j]) return -2; /* "Interaction between residues of different chains" */
        double sig2 = pow(abs(i - j), 2 * c->fm_sigma_exp);
        sig2 = sig2 * c->ff_ww * c->ff_ww;
        double g = gm_get2(&c->fmg, itype, jtype, i, j);
        if (c->ff_seqsep) g *= pow((abs(i - j) - minsep + 1), -c->ff_seqsep_gamma);
        double ekwg = ekw * g;
        const double *xi[4] = {XCA(i), XCA(i), XCB(i), XCB(i)}, *xj[4] = {XCA(j), XCB(j), XCA(j), XCB(j)};
        for (int k = 0; k < 4; ++k) {
          if (ity[k] == FM_CB && (c->se[i] == 'G' || m->se[i - m->pos] == 'G')) continue;
          if (jty[k] == FM_CB && (c->se[j] == 'G' || m->se[j - m->pos] == 'G')) continue;
          double r;
          if (dec) {
            r = mem_rf(dec, i, ity[k], j, jty[k]);
            if (isnan(r)) r = 0.0;
          } else {
            double dx[3] = {xi[k][0] - xj[k][0], xi[k][1] - xj[k][1], xi[k][2] - xj[k][2]};
            r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
          }
          double rf = mem_rf(m, i, ity[k], j, jty[k]);
          double dr = r - rf, drsq = dr * dr;
          /* the reference's normalizeInteraction scales V before assigning
             it (fb.cpp:4548-4553): no effect */
          double V = -ekwg * exp(-drsq / (2 * sig2));
          E[i] += V;
          E[j] += V;
        }
      }
    }
  }
  return 0;
}

/* one fragment-frustratometer calculation at step `step` (fb.cpp:7537-
 * 7576) at the given coordinates.  rows: n x 3 (frustration index
 * (native - mean) / std, native - mean, std, per residue).  gen (read mode,
 * may be NULL): at the calculation that computes the decoy statistics, the
 * n native energies then the n x n_decoys decoy energies (residue-major),
 * the values written to fragment_frustration_native.dat / _decoy.dat
 * (4786-4802); *gen_written is 1 then.  Returns 0, or < 0 on the
 * reference's errors. */
int or_frag_frust(or_ctx *c, const double *xca, const double *xcb, const double *xo, long long step, double *rows,
                  double *gen, int *gen_written)
{
  int n = c->n, nc = c->ff_ncalc;
  c->xca = xca; c->xcb = xcb; c->xo = xo;
  *gen_written = 0;
  double *E = (double *)calloc((size_t)n * nc, sizeof(double)); /* E[calc * n + i] */
  int rc = 0;
  if (c->ff_mode == 0) {
    for (int d = 0; d < nc && !rc; ++d) {
      rc = d == 0 ? ff_energy(c, c->mems, c->frag_mem_map, c->ilen_fm_map, NULL, E)
                  : ff_energy(c, c->ff_mems, c->ff_map, c->ff_ilen, NULL, E + (size_t)d * n);
      /* randomize_decoys (fb.cpp:4565-4611) */
      for (int k = 0; k < c->ff_nm; ++k) c->ff_mems[k].pos = or_rng_next(c) % (n - c->ff_mems[k].len + 1);
      if (ff_build_map(c)) rc = -3;
    }
    if (!rc) {
      /* compute_fragment_frustration, fb.cpp:4613-4689 */
      for (int i = 0; i < n; ++i) {
        E[i] /= c->n_mems;
        for (int d = 1; d < nc; ++d) E[(size_t)d * n + i] /= c->ff_nm;
      }
      for (int i = 0; i < n; ++i) {
        double mean = 0.0, var = 0.0;
        for (int d = 1; d < nc; ++d) mean += E[(size_t)d * n + i];
        mean /= (nc - 1);
        for (int d = 1; d < nc; ++d) var += pow(E[(size_t)d * n + i] - mean, 2);
        var /= (nc - 1);
        rows[3 * i] = (E[i] - mean) / sqrt(var);
        rows[3 * i + 1] = E[i] - mean;
        rows[3 * i + 2] = sqrt(var);
      }
    }
  } else {
    rc = ff_energy(c, c->mems, c->frag_mem_map, c->ilen_fm_map, NULL, E);
    /* the decoy statistics: at step 0 in the reference (fb.cpp:7566-7568;
       left unset when a run starts elsewhere), here at the first call */
    if (!rc && !c->ff_gen_done) {
      int nd = c->ff_nm;
      for (int d = 0; d < nd && !rc; ++d) rc = ff_energy(c, c->mems, c->frag_mem_map, c->ilen_fm_map, &c->ff_mems[d], E + (size_t)(d + 1) * n);
      for (int i = 0; i < n && !rc; ++i) {
        double mean = 0.0, var = 0.0;
        for (int d = 1; d < nc; ++d) mean += E[(size_t)d * n + i];
        mean /= nd;
        for (int d = 1; d < nc; ++d) var += pow(E[(size_t)d * n + i] - mean, 2);
        var /= nd;
        c->ff_mean[i] = mean;
        c->ff_var[i] = var;
      }
      if (!rc && gen) {
        for (int i = 0; i < n; ++i) gen[i] = E[i];
        for (int i = 0; i < n; ++i)
          for (int d = 0; d < nd; ++d) gen[n + (size_t)i * nd + d] = E[(size_t)(d + 1) * n + i];
        *gen_written = 1;
      }
      c->ff_gen_done = 1;
    }
    if (!rc)
      for (int i = 0; i < n; ++i) {
        rows[3 * i] = (E[i] - c->ff_mean[i]) / sqrt(c->ff_var[i]);
        rows[3 * i + 1] = E[i] - c->ff_mean[i];
        rows[3 * i + 2] = sqrt(c->ff_var[i]);
      }
  }
  free(E);
  (void)step;
  return rc;
}

int or_frag_frust_dims(const or_ctx *c, int *mode, int *ncalc, int *ndecoys)
{
  *mode = c->frag_frust_flag ? c->ff_mode : -1;
  *ncalc = c->ff_ncalc;
  *ndecoys = c->ff_nm;
  return 0;
}
