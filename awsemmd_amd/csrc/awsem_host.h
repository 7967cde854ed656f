// awsem_host.h -- host-side system description: the force-field parameters
// and one system's sequence, chains, ssweight and fragment memories, as the
// FixBackbone constructor builds them (src/fix_backbone.cpp:86-1072).
#pragma once
#include <cmath>
#include <map>
#include <string>
#include <cstring>
#include <limits>
#include <vector>
#include "awsem_types.h"

struct AwStatus {
  int code = 0;
  std::string msg;
  bool ok() const { return code == 0; }
  static AwStatus err(int c, const std::string &m) { AwStatus s; s.code = c; s.msg = m; return s; }
};

// Gamma_Array (src/fragment_memory.cpp:275-519)
struct AwGamma {
  int nseq_cl = 0, nres_cl = 0;
  bool frag_resty = false;
  int i_sep[10] = {0};
  std::vector<double> gamma;
  int min_sep() const { return i_sep[0]; }
  int max_sep() const { return i_sep[nseq_cl]; }
  double get(int it, int jt, int ires, int jres) const;
};

struct AwMemory {
  int pos = 0, len = 0;
  double weight = 1.0;
  std::string se;               // one-letter names from the .gro CA records
  std::vector<double> ca, cb;   // len*3, Angstrom
};

// fragment-memory table mode: curve ids per (i, j - i - minSep, atom pair
// k) at 4 (tb_nbrs i + j - i - minSep) + k: -1 no curve, -2 a curve of
// zeros (read from a table file), else the index of its tb_size energies
// and forces in e / f
struct AwFMTable {
  int tb_size = 0, tb_nbrs = 0;
  std::vector<int> cid;
  std::vector<double> e, f;
  int n_curves() const { return tb_size ? (int)(e.size() / tb_size) : 0; }
};

// coeff-toggled terms' side files (fb.cpp:607-620, 695-743, 1018-1046,
// 1106-1170)
struct AwExtras {
  std::vector<int> zres;              // membrane class per residue (zim)
  std::vector<double> charge;         // Debye-Hueckel charge per residue
  // contact restraints: entries sorted by (lower, upper residue, file order)
  std::vector<int> cr_i, cr_j;
  std::vector<double> cr_w, cr_r0;    // w already times k_cont_rest
  // AMH-Go: gamma classes, the native structure (one memory over the
  // system), censored pairs (flag 1, n*n) or native distances (flag 2,
  // CA-CA, CB-CB, CA-CB, n*n each), normalisation
  AwGamma amhg;
  AwMemory amh_mem;
  std::vector<unsigned char> censor;
  std::vector<double> rn[3];
  double go_norm = 1.0;
};

struct AwSystem {
  FFParams P;
  int n = 0, nch = 0;
  std::string se;
  std::vector<int> ch_start, ch_len;   // 0-based start
  std::vector<double> aps;             // n*12
  std::vector<int> rinfo;
  AwGamma fmg;
  std::vector<AwMemory> mems;
  AwFMTable fmt;                       // table mode only
  AwExtras xt;                         // coeff-toggled terms
  std::string cr_file;                 // contact-restraints file named by the coeff
  std::string ff_file;                 // the fragment frustratometer's decoy library
  std::vector<AwMemory> ff_mems;       // its fragments (positions as read) or structures
  char frag_mems_file[256] = {0}, fm_gamma_file[256] = {0};
};

int aw_restype(char c);
AwStatus aw_parse_coeff(const std::string &path, FFParams &P, std::string &mem_file, std::string &gamma_file,
                        std::string &cr_file, std::string &ff_file);
// the fragment frustratometer's decoy library (after the memories)
AwStatus aw_read_decoy_mems(const std::string &dir, AwSystem &S);
// the coeff-toggled terms' side files (after the sequence, tables and
// memories): contact list, zim, charges, AMH-Go files; their share of
// pair_list_cutoff (fb.cpp:1058-1066) and the AMH-Go normalisation
AwStatus aw_read_extras(const std::string &dir, const std::string &cr_file, AwSystem &S);
// AMH-Go native distance of (i, ia) - (j, ja), ia/ja 0 CA, 1 CB
double aw_go_rnative(const AwSystem &S, int i, int ia, int j, int ja);
// the device items of the coeff-toggled terms (AwExtras + memories)
struct AwXtPack {
  std::vector<int> go_start;          // [2n + 1], empty without AMH-Go
  std::vector<XtGo> go;
  std::vector<XtPair> cr, dh, vfm;
};
AwStatus aw_build_extras(const AwSystem &S, const FFParams &P, AwXtPack &X);
bool aw_same_xt_params(const FFParams &a, const FFParams &b);
AwStatus aw_read_param_tables(const std::string &dir, AwSystem &S);
AwStatus aw_read_sequence(const std::string &path, AwSystem &S);
AwStatus aw_read_ssweight(const std::string &dir, AwSystem &S);
AwStatus aw_read_gamma_array(const std::string &path, AwGamma &g);
AwStatus aw_read_mems(const std::string &dir, const std::string &mem_file, AwSystem &S);
AwStatus aw_finish_system(AwSystem &S);

struct AwFMPack {
  std::vector<double> coef;   // epsilon * k_frag_mem * weight per memory
  std::vector<double> rf;     // 4 per (memory, pair): CA-CA, CA(i)-CB(j), CA(j)-CB(i), CB-CB; -1 = no atom
  std::vector<double> items;  // gamma, log2(e)/(2 sigma^2), 1/sigma^2 per item
  std::vector<int> meta;      // ij, rf0, stride, nmem, coef0 per item
  int n_items() const { return (int)(items.size() / 3); }
};
// appends one system's FM items (residue indices local to the system)
// P: the force field (fragment-memory strength, sigma exponent, epsilon)
AwStatus aw_build_fm(const AwSystem &S, const FFParams &P, AwFMPack &B);
// table mode: the curves from the memories (fb.cpp:4238-4323), or read
// from fm_table.energy / fm_table.force (fb.cpp:4179-4236), and written
// there (fb.cpp:6062-6106)
AwStatus aw_build_fm_table(AwSystem &S);
AwStatus aw_read_fm_table(const std::string &dir, AwSystem &S);
AwStatus aw_write_fm_table(const std::string &dir, const AwSystem &S);
// the pair path's table items (fb.cpp:7513-7527), one per atom pair that
// table_fragment_memory evaluates (fb.cpp:4325-4429), 4 ints each: i, j,
// the atom pair q (0 CA-CA, 1 CA(i)-CB(j), 2 CB(i)-CA(j), 3 CB-CB) and its
// curve id, -1 for a range check only (a curve of zeros, or a missing curve:
// then the pair's later atom pairs are not listed, the reference returns
// there).  Pairs whose CA-CA curve is missing, and atom pairs skipped by a
// target glycine CB, have no item.
void aw_fm_table_items(const AwSystem &S, std::vector<int> &items);
// the device form of a pack in precision T: work items with the width
// factor sc = sqrt(c2) rounded to T, and the Rf entries pre-multiplied by
// it (in T, as the kernels scale the distances), skipped entries (memory or
// target glycine CB) set to AW_FM_FAR.  Every Rf entry belongs to exactly
// one item.
// Weights: a group whose weights are all positive gets log2(w) (padding
// -inf) and AW_FM_LOGW on its items, so that w exp2(-d^2) = exp2(log2 w - d^2)
// costs no multiply; other groups keep w.
template <class T, class ItemT>
void aw_fm_device(const AwFMPack &B, std::vector<ItemT> &items, std::vector<T> &rf, std::vector<T> &coef)
{
  items.assign(B.n_items(), ItemT());
  rf.assign(B.rf.size(), (T)AW_FM_FAR);
  coef.assign(B.coef.begin(), B.coef.end());
  std::vector<signed char> logw(B.coef.size(), -1);  // per group start: -1 unseen, 0 no, 1 yes
  for (int k = 0; k < B.n_items(); ++k) {
    const int *m = &B.meta[5 * k];
    int c0 = m[4], nm = m[3] & AW_FM_NMEM_MASK;
    if (logw[c0] < 0) {
      bool pos = true;
      for (int q = 0; q < nm; ++q) pos = pos && B.coef[c0 + q] > 0.0;
      logw[c0] = pos ? 1 : 0;
      if (pos) {
        int len = (nm + AW_FM_PAD - 1) / AW_FM_PAD * AW_FM_PAD;
        for (int q = 0; q < len; ++q)
          coef[c0 + q] = q < nm ? (T)std::log2(B.coef[c0 + q]) : -std::numeric_limits<T>::infinity();
      }
    }
  }
  for (int k = 0; k < B.n_items(); ++k) {
    ItemT &it = items[k];
    std::memset(&it, 0, sizeof it);
    const int *m = &B.meta[5 * k];
    T sc = (T)std::sqrt(B.items[3 * k + 1]);
    it.gamma = (T)B.items[3 * k];
    it.sc = sc;
    it.invs = (T)(B.items[3 * k + 2] / (double)sc);
    it.ij = m[0]; it.rf0 = m[1]; it.stride = m[2]; it.nmem = m[3] | (logw[m[4]] > 0 ? AW_FM_LOGW : 0);
    it.coef0 = m[4];
    bool gi = m[3] & AW_FM_GLY_I, gj = m[3] & AW_FM_GLY_J;
    bool live[4] = {true, !gj, !gi, !gi && !gj};
    int nmem = m[3] & AW_FM_NMEM_MASK;
    for (int q = 0; q < nmem; ++q) {
      size_t e = 4 * ((size_t)m[1] + (size_t)q * (size_t)m[2]);
      for (int c = 0; c < 4; ++c) {
        double v = B.rf[e + c];
        if (live[c] && v >= 0.0) rf[e + c] = (T)v * sc;
      }
    }
  }
}
void aw_default_pair_styles(FFParams &P);
double aw_ssb_cut(const FFParams &P);
// ---------------------------------------------------------------------------
// One system's device image (awsem_devsys.cpp): a single allocation holding
// its per-residue constants, fragment-memory tables, Verlet list block and,
// optionally, its state (x, v, f, energy) and a global working slab.
struct AwDevPlan {
  AwFMPack fm;
  bool fp32 = false, scratch = false, own_state = false;
  int nl_cap = 0, n = 0, nch = 0, n_items = 0;
  size_t off_x = 0, off_v = 0, off_f = 0, off_e = 0, off_rinfo = 0, off_aps = 0, off_chs = 0, off_chl = 0;
  size_t off_items = 0, off_rf = 0, off_coef = 0, off_nl = 0, off_nlx = 0, off_scr = 0, off_cells = 0;
  size_t off_tbi = 0, off_tbc = 0, off_fault = 0;  // table items, curves, error word
  int n_tb_items = 0;
  // coeff-toggled terms: AMH-Go start offsets and pairs, contact restraints,
  // Debye-Hueckel pairs, vector-memory pairs, membrane classes
  size_t off_gos = 0, off_go = 0, off_cr = 0, off_dh = 0, off_vfm = 0, off_zres = 0;
  int n_go = 0, n_cr = 0, n_dh = 0, n_vfm = 0;
  bool has_go = false, has_zres = false;
  double go_norm = 1.0, list_r2 = 0.0, cr_cut2 = 0.0;
  int cell_h = 0;
  std::vector<int> fm_ij;  // (i << 16) | j of every fragment-memory item, in device order
  size_t bytes = 0;
  std::vector<unsigned char> image;  // host copy of the whole device image
};
struct AwDevSystem {
  void *arena = nullptr;
  size_t bytes = 0;
  ReplicaDesc d;  // device pointers into the arena; schedule fields unset
  AwDevSystem() { std::memset(&d, 0, sizeof d); }
  ~AwDevSystem();
  AwDevSystem(const AwDevSystem &) = delete;
  AwDevSystem &operator=(const AwDevSystem &) = delete;
};
// P: the force field to pack with (S.P when null); lib = false leaves the
// fragment-memory library (items, distances, coefficients, table items and
// curves) out of the image: the caller points the descriptor at another
// replica's copy of the same system's library
AwStatus aw_devsys_plan(const AwSystem &S, const FFParams *P, bool fp32, bool scratch, bool own_state, AwDevPlan &plan,
                        bool lib = true);
// allocates the device image and uploads plan.image (the plan's host copy)
AwStatus aw_devsys_create(const AwDevPlan &plan, AwDevSystem &out);
AwStatus upload_raw(void **dst, const void *src, size_t bytes);
long long aw_slab_doubles(int n);


// ---------------------------------------------------------------------------
// the tertiary frustratometer of a context (awsem_frust.hip): glibc's rand()
// stream the decoys draw from (the reference's unseeded rand(), continued
// across calls), the configurational decoy statistics (computed once,
// fb.cpp:5099, 5342) and the output files
struct AwFrustState {
  uint32_t rng[34];
  int rng_k = 0;  // ring slot of the next value
  bool done = false;
  double stats[2] = {0.0, 0.0};
  FILE *dat = nullptr, *tcl = nullptr;
  // the n-mer frustratometer's files and the trap drawing's representation
  // counter (a function-static in the reference, fb.cpp:5802)
  FILE *nm_dat = nullptr, *nm_tcl = nullptr, *nm_trap = nullptr;
  int nm_rep = 1;
  // the fragment frustratometer: its five files, the decoy fragments'
  // current positions (shuffle), read mode's per-residue decoy statistics
  FILE *ff_out[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  std::vector<int> ff_pos;
  bool ff_gen_done = false;
  std::vector<double> ff_mean, ff_var;
  // its call-invariant device data, built at the first call: the gamma
  // table; the memories' items, positions and weights; the decoy
  // fragments' items and weights
  void *ff_dev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int ff_n_nat = 0, ff_n_dec = 0;
  bool ff_ready = false;
  AwFrustState();
  ~AwFrustState();
  int next();
  void seed(unsigned s);
};
struct awsem_ctx;
// the context's device image built and xres (residue layout) uploaded as
// its coordinates: FFParams, rinfo and coordinates on the device, stream
struct AwCtxDev {
  const FFParams *params;
  const int *rinfo;
  const double *x;
  void *stream;
};
AwStatus aw_ctx_device(awsem_ctx *c, const double *xres, AwCtxDev &out);
AwFrustState &aw_ctx_frust(awsem_ctx *c);
const AwSystem *aw_ctx_system(const awsem_ctx *c);
const std::string &aw_ctx_dir(const awsem_ctx *c);
int aw_ctx_fail(awsem_ctx *c, const AwStatus &s);
