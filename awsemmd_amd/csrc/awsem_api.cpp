// awsem_api.cpp -- the C ABI of include/awsem.h over the host parsers and
// the gfx950 kernels.
#include "../../include/awsem.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <unordered_map>
#include <string>
#include <vector>

#include "awsem_host.h"
#include "awsem_launch.h"

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return AwStatus::err(AWSEM_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// device copy of a set of systems (the static batch, and the single-system
// context as a batch of one)

struct DevBatch {
  int nrep = 0, max_n = 0, total_res = 0, fp32 = 0, lds = 0;
  std::vector<std::unique_ptr<AwDevSystem>> sys;
  std::vector<ReplicaDesc> reps;   // host copy; schedule fields set per launch
  std::vector<int> res_off;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  ReplicaDesc *d_rep = nullptr;
  void *d_params = nullptr;
  int *d_queue = nullptr, *d_order = nullptr, *d_count = nullptr;
  int *d_fault = nullptr;  // error word of every replica (table mode: AW_FAULT_TB_RANGE)
  bool has_table = false;  // some replica evaluates table pairs
  double *d_x = nullptr, *d_v = nullptr, *d_f = nullptr, *d_e = nullptr;  // concatenated state
  long long *d_list = nullptr;
  unsigned long long *d_prof = nullptr;  // phase clock (AWSEM_PHASE_PROF=1)
  bool nl_dirty = true;
  bool f_valid = false;  // d_f holds the forces of d_x (after a completed launch)
  int list_cap = 0;
  FFParams P;
  double last_ms = 0.0;
  // the resident launch sequence captured as a graph, replayed per call;
  // recaptured when the kernel arguments (mode, dt, damp) change
  bool use_graph = true;
  hipGraphExec_t graph = nullptr;
  int g_mode = -1, g_engine = -1, g_nsteps = -1;
  double g_dt = 0.0, g_damp = 0.0;
  int engine = 0;  // 0 resident (workgroup per replica), 1 wide (phase launches over all replicas)
  int last_engine = -1;  // engine of the last completed launch (where its queue lengths are kept)
  // fragment-memory sharing of the resident engine: one allocation holding
  // the claim words, the list of sharing replicas and their chunk tables,
  // position copies and accumulators
  void *d_share = nullptr;
  std::vector<unsigned char> share_img;  // its initial contents (restored after a failed launch)
  int nshare = 0, share_spread = 1;
  int *d_share_rep = nullptr, *d_share_n = nullptr;
  unsigned long long *d_share_word = nullptr;

  ~DevBatch()
  {
    if (graph) (void)hipGraphExecDestroy(graph);
    void *ptrs[] = {d_rep, d_params, d_queue, d_order, d_count, d_x, d_v, d_f, d_e, d_list, d_prof, d_share, d_fault};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    sys.clear();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  AwBatchPtrs ptrs() const
  {
    AwBatchPtrs p;
    p.rep = d_rep; p.nrep = nrep; p.max_n = max_n; p.fp32 = fp32; p.lds_resident = lds; p.table = P.fm_tb_flag;
    p.extras = P.xt_on;
    p.params = d_params; p.queue = d_queue; p.order = d_order; p.prof = d_prof;
    p.nshare = nshare; p.share_rep = d_share_rep; p.share_n = d_share_n; p.share_word = d_share_word;
    p.share_spread = share_spread;
    return p;
  }

  // schedule of every replica for one launch, then the descriptors to the device
  AwStatus set_schedule(int nsteps, const awsem_langevin *lg)
  {
    for (int r = 0; r < nrep; ++r) {
      ReplicaDesc &R = reps[r];
      R.nsteps = nsteps; R.ramp0 = 0; R.run_steps = nsteps; R.rng_id = r;
      R.t_start = lg ? lg->t_start : 0.0; R.t_stop = lg ? lg->t_stop : 0.0;
      R.seed = lg ? lg->seed : 0; R.step0 = lg ? lg->step0 : 0;
      R.flags = (f_valid ? AW_RF_F_VALID : 0) | (nl_dirty ? AW_RF_NL_RESET : 0);
    }
    HIPCHK(hipMemcpyAsync(d_rep, reps.data(), reps.size() * sizeof(ReplicaDesc), hipMemcpyHostToDevice, stream));
    return AwStatus();
  }
};

static AwStatus launch_seq(DevBatch &D, const AwRun &run, int engine, int nst);

// the batch's error word (every replica's kernels raise their bits in it;
// table mode: a distance outside the table's range), read after the
// launch's synchronisation and cleared when set
static AwStatus take_faults(DevBatch &D)
{
  if (!D.has_table) return AwStatus();
  int any = 0;
  if (hipMemcpy(&any, D.d_fault, sizeof any, hipMemcpyDeviceToHost) != hipSuccess)
    return AwStatus::err(AWSEM_E_DEVICE, "reading the error word");
  if (any) {
    int z = 0;
    if (hipMemcpy(D.d_fault, &z, sizeof z, hipMemcpyHostToDevice) != hipSuccess)
      return AwStatus::err(AWSEM_E_DEVICE, "clearing the error word");
  }
  if (any & AW_FAULT_TB_RANGE) return AwStatus::err(AWSEM_E_RANGE, "Table Fragment Memory: r is out of computed range.");
  return AwStatus();
}

template <class T>
static AwStatus upload(T **dst, const std::vector<T> &src, size_t min_elems = 1)
{
  size_t n = std::max(src.size(), min_elems);
  HIPCHK(hipMalloc((void **)dst, n * sizeof(T)));
  if (!src.empty()) HIPCHK(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return AwStatus();
}

// fragment-memory chunks of one replica: consecutive items, at most cs of
// them, whose residues lie in [lo, lo + AW_SHARE_SPAN); false when the
// items do not cut into AW_SHARE_MAXCH such chunks
static bool fm_chunks(const std::vector<int> &ij, int cs, std::vector<ShareChunk> &out)
{
  out.clear();
  int nit = (int)ij.size();
  int k0 = 0, lo = 1 << 30, hi = -1;
  for (int k = 0; k <= nit; ++k) {
    int i = k < nit ? ij[k] >> 16 : 0, j = k < nit ? ij[k] & 0xffff : 0;
    bool close = k == nit || k - k0 == cs ||
                 std::max(hi, j) - std::min(lo, i) + 1 > AW_SHARE_SPAN;
    if (close && k > k0) {
      ShareChunk c;
      c.k0 = k0; c.k1 = k; c.lo = lo; c.pad = 0;
      out.push_back(c);
      k0 = k; lo = 1 << 30; hi = -1;
    }
    if (k == nit) break;
    if (j - i + 1 > AW_SHARE_SPAN) return false;
    lo = std::min(lo, i);
    hi = std::max(hi, j);
  }
  return (int)out.size() <= AW_SHARE_MAXCH;
}

// fragment-memory sharing for a resident batch in LDS: replicas above the
// batch's mean size (AWSEM_SHARE_MIN_N overrides; AWSEM_SHARE=0 turns it
// off) hand chunks of AWSEM_SHARE_CHUNK (512) items to idle workgroups
static AwStatus setup_share(DevBatch &D, const std::vector<std::vector<int>> &ij)
{
  const char *on = std::getenv("AWSEM_SHARE");
  if ((on && std::atoi(on) == 0) || !D.lds || D.nrep < 2) return AwStatus();
  // a helping workgroup's waves stage residue spans in the slab
  if (aw_slab_doubles(D.max_n) < 8LL * 12 * AW_SHARE_SPAN) return AwStatus();
  double min_n = (double)D.total_res / D.nrep;
  if (const char *e = std::getenv("AWSEM_SHARE_MIN_N")) min_n = std::atof(e);
  int cs = 512;
  if (const char *e = std::getenv("AWSEM_SHARE_CHUNK")) cs = std::max(64, std::atoi(e));
  if (const char *e = std::getenv("AWSEM_SHARE_SPREAD")) {
    // helpers take the first open replica, longest first, from the
    // (blockIdx mod spread)-th on: 1 (the default) goes strictly to the
    // longest open replica, which measured faster than spreading helpers
    // over the 64 longest (2.31 M against 2.22 M replica-steps/s)
    int v = std::atoi(e), p2 = 1;
    while (p2 * 2 <= v && p2 < 64) p2 *= 2;
    D.share_spread = p2;
  }
  std::vector<int> rep;
  std::vector<std::vector<ShareChunk>> ch;
  for (int r = 0; r < D.nrep; ++r) {
    if (D.reps[r].n < min_n || ij[r].empty()) continue;
    std::vector<ShareChunk> c;
    if (!fm_chunks(ij[r], cs, c) || c.size() < 2) continue;
    rep.push_back(r);
    ch.push_back(std::move(c));
  }
  if (rep.empty()) return AwStatus();
  // longest first
  std::vector<int> ord(rep.size());
  for (size_t q = 0; q < ord.size(); ++q) ord[q] = (int)q;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return D.reps[rep[a]].n > D.reps[rep[b]].n; });
  int ns = (int)rep.size();
  // layout: words, done counters, replica list, chunk counts, then per
  // replica its chunks, positions and accumulator
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t o_word = 0, o_done = al(8 * (size_t)ns), o_rep = o_done + al(4 * (size_t)ns), o_n = o_rep + al(4 * (size_t)ns);
  size_t cur = o_n + al(4 * (size_t)ns);
  std::vector<size_t> o_ch(ns), o_x(ns), o_acc(ns);
  for (int q = 0; q < ns; ++q) {
    int r = rep[ord[q]], n = D.reps[r].n;
    o_ch[q] = cur; cur += al(ch[ord[q]].size() * sizeof(ShareChunk));
    o_x[q] = cur; cur += al(6 * (size_t)n * sizeof(double));
    o_acc[q] = cur; cur += al((6 * (size_t)n + 1) * sizeof(double));
  }
  std::vector<unsigned char> img(cur, 0);
  HIPCHK(hipMalloc(&D.d_share, cur));
  unsigned char *b = (unsigned char *)D.d_share;
  for (int q = 0; q < ns; ++q) {
    int r = rep[ord[q]];
    const std::vector<ShareChunk> &c = ch[ord[q]];
    unsigned long long w = c.size();  // closed: no chunk open before the owner's first evaluation
    std::memcpy(img.data() + o_word + 8 * q, &w, 8);
    std::memcpy(img.data() + o_rep + 4 * q, &r, 4);
    int nc = (int)c.size();
    std::memcpy(img.data() + o_n + 4 * q, &nc, 4);
    std::memcpy(img.data() + o_ch[q], c.data(), c.size() * sizeof(ShareChunk));
    ReplicaDesc &R = D.reps[r];
    R.sh_word = (unsigned long long *)(b + o_word) + q;
    R.sh_done = (int *)(b + o_done) + q;
    R.sh_chunks = (const ShareChunk *)(b + o_ch[q]);
    R.sh_x = (double *)(b + o_x[q]);
    R.sh_acc = (double *)(b + o_acc[q]);
    R.sh_n = nc;
    R.sh_q = q;
  }
  HIPCHK(hipMemcpy(D.d_share, img.data(), cur, hipMemcpyHostToDevice));
  D.share_img = std::move(img);
  D.nshare = ns;
  D.d_share_word = (unsigned long long *)(b + o_word);
  D.d_share_rep = (int *)(b + o_rep);
  D.d_share_n = (int *)(b + o_n);
  return AwStatus();
}

// two 64-bit hashes of a byte range (8-byte words; FNV-style multiply-xor
// with different primes and seeds): the key under which identical
// fragment-memory libraries are found
static void bytes_hash2(const unsigned char *p, size_t nb, uint64_t &h1, uint64_t &h2)
{
  h1 = 1469598103934665603ULL;
  h2 = 0x9E3779B97F4A7C15ULL ^ nb;
  size_t k = 0;
  for (; k + 8 <= nb; k += 8) {
    uint64_t w;
    std::memcpy(&w, p + k, 8);
    h1 = (h1 ^ w) * 1099511628211ULL;
    h2 = (h2 ^ (w + (h2 >> 29))) * 0xFF51AFD7ED558CCDULL;
  }
  for (; k < nb; ++k) {
    h1 = (h1 ^ p[k]) * 1099511628211ULL;
    h2 = (h2 ^ (p[k] + (h2 >> 29))) * 0xFF51AFD7ED558CCDULL;
  }
}

static AwStatus make_dev_batch(const std::vector<const AwSystem *> &sys, int precision, bool want_lds,
                               std::unique_ptr<DevBatch> &out)
{
  auto D = std::unique_ptr<DevBatch>(new DevBatch());
  D->nrep = (int)sys.size();
  D->fp32 = precision == AWSEM_PREC_FP32;
  D->P = sys[0]->P;
  for (const AwSystem *S : sys)
    if (!aw_same_xt_params(S->P, D->P))
      return AwStatus::err(AWSEM_E_ARG, "batch replicas must share one coeff parameter set (the optional terms differ)");
  int res = 0;
  for (const AwSystem *S : sys) { D->res_off.push_back(res); res += S->n; D->max_n = std::max(D->max_n, S->n); }
  D->total_res = res;
  D->lds = want_lds && D->max_n <= aw_lds_max_residues();
  aw_resident_prepare();
  HIPCHK(hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&D->ev0));
  HIPCHK(hipEventCreate(&D->ev1));
  size_t st = std::max<size_t>((size_t)res * 9, 1) * sizeof(double);
  HIPCHK(hipMalloc((void **)&D->d_x, st));
  HIPCHK(hipMalloc((void **)&D->d_v, st));
  HIPCHK(hipMalloc((void **)&D->d_f, st));
  HIPCHK(hipMemset(D->d_x, 0, st));
  HIPCHK(hipMemset(D->d_v, 0, st));
  HIPCHK(hipMemset(D->d_f, 0, st));
  HIPCHK(hipMalloc((void **)&D->d_e, (size_t)D->nrep * AW_N_ET * sizeof(double)));
  HIPCHK(hipMemset(D->d_e, 0, (size_t)D->nrep * AW_N_ET * sizeof(double)));
  HIPCHK(hipMalloc((void **)&D->d_fault, 256));
  HIPCHK(hipMemset(D->d_fault, 0, 256));
  std::vector<std::vector<int>> fm_ij(D->nrep);
  // replicas of one system (the same fragment-memory library, byte for
  // byte) read a single copy of its tables: one stream of them through the
  // caches instead of one per replica
  // (a context given more than once is found by address, and its later
  // replicas carry no library of their own; separate contexts with equal
  // libraries by a hash of the packed bytes)
  struct FmLib { size_t bytes; uint64_t h1, h2; int n, rep; };
  std::vector<FmLib> libs;
  std::unordered_map<const AwSystem *, int> first_of;
  for (int r = 0; r < D->nrep; ++r) {
    const AwSystem &S = *sys[r];
    auto seen = first_of.find(&S);
    int same = seen == first_of.end() ? -1 : seen->second;
    if (same < 0) first_of[&S] = r;
    AwDevPlan plan;
    // a global working slab for every replica: the wide engine, and
    // replicas too large for LDS, work from it
    AwStatus s = aw_devsys_plan(S, nullptr, D->fp32, want_lds || !D->lds, false, plan, same < 0);
    if (!s.ok()) return s;
    int src = same;
    FmLib lib{plan.off_nl - plan.off_items, 0, 0, S.n, r};
    if (same < 0 && plan.n_items + plan.n_tb_items > 0) {
      bytes_hash2(plan.image.data() + plan.off_items, lib.bytes, lib.h1, lib.h2);
      for (const FmLib &L : libs)
        if (L.bytes == lib.bytes && L.h1 == lib.h1 && L.h2 == lib.h2 && L.n == lib.n) { src = L.rep; break; }
      if (src < 0) libs.push_back(lib);
    }
    std::unique_ptr<AwDevSystem> ds(new AwDevSystem());
    if (!(s = aw_devsys_create(plan, *ds)).ok()) return s;
    ReplicaDesc R = ds->d;
    if (src >= 0) {
      const ReplicaDesc &Q = D->reps[src];
      R.items = Q.items; R.rf = Q.rf; R.coef = Q.coef; R.fm_items = Q.fm_items;
      R.tb_item = Q.tb_item; R.tb_curve = Q.tb_curve; R.tb_items = Q.tb_items;
      if (same >= 0 && want_lds) plan.fm_ij = fm_ij[src];
    }
    R.x = D->d_x + 9LL * D->res_off[r];
    R.v = D->d_v + 9LL * D->res_off[r];
    R.f = D->d_f + 9LL * D->res_off[r];
    R.energy = D->d_e + (long long)r * AW_N_ET;
    R.fault = D->d_fault;
    D->has_table = D->has_table || R.tb_items > 0;
    D->reps.push_back(R);
    D->sys.push_back(std::move(ds));
    if (want_lds) fm_ij[r] = std::move(plan.fm_ij);
  }
  AwStatus s;
  if (want_lds && !(s = setup_share(*D, fm_ij)).ok()) return s;
  if (!(s = upload_raw((void **)&D->d_rep, D->reps.data(), D->reps.size() * sizeof(ReplicaDesc))).ok()) return s;
  if (!(s = upload_raw(&D->d_params, &D->P, sizeof(FFParams))).ok()) return s;
  // [0] work queue, [16] finished replicas, [32] sharing error flag
  HIPCHK(hipMalloc((void **)&D->d_queue, 256));
  HIPCHK(hipMemset(D->d_queue, 0, 256));
  HIPCHK(hipMalloc((void **)&D->d_count, 64));
  if (const char *e = std::getenv("AWSEM_PHASE_PROF")) {
    if (std::atoi(e) > 0) {
      HIPCHK(hipMalloc((void **)&D->d_prof, 32 * sizeof(unsigned long long)));
      HIPCHK(hipMemset(D->d_prof, 0, 32 * sizeof(unsigned long long)));
    }
  }
  // longest-first order for the resident engine's work queue
  std::vector<int> order(D->nrep);
  for (int r = 0; r < D->nrep; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return D->reps[a].n > D->reps[b].n; });
  if (!(s = upload(&D->d_order, order)).ok()) return s;
  HIPCHK(hipDeviceSynchronize());
  out = std::move(D);
  return AwStatus();
}

// ---------------------------------------------------------------------------

struct awsem_ctx {
  AwSystem sys;
  int precision = AWSEM_PREC_FP64;
  std::string err;
  bool topo = false;
  int n_atoms = 0;
  std::vector<int> atom_of;   // [3*n]: atom index of (kind, residue)
  std::unique_ptr<DevBatch> dev;
  bool dirty = true;          // parameters changed since dev was built
  // periodic box (domain->x/y/zperiodic and prd): coordinates are unwrapped
  // with the atoms' image flags before the evaluation (fb.cpp:7067-7112)
  int periodic[3] = {0, 0, 0};
  double prd[3] = {0.0, 0.0, 0.0};
  double last_e[AW_N_ET] = {0.0};  // energies of the last evaluation
  FILE *elog = nullptr;            // energy.log (awsem_energy_log_open)
  std::string dir = ".";           // working directory of the side and output files
  AwFrustState tf;                 // the tertiary frustratometer's state
  ~awsem_ctx()
  {
    if (elog) std::fclose(elog);
  }
};

struct awsem_batch {
  std::unique_ptr<DevBatch> dev;
  std::string err;
  double ms_total = 0.0;
  int n_kernels = 0;
};

const AwSystem *aw_ctx_system(const awsem_ctx *c) { return c ? &c->sys : nullptr; }

static int fail(awsem_ctx *c, const AwStatus &s)
{
  if (c) c->err = s.msg;
  return s.code;
}

// table mode: fm_use_table set and both table files present (fb.cpp:746)
static bool fm_tables_stored(const FFParams &P, const std::string &dir)
{
  if (!P.fm_tb_flag || !P.fm_use_table) return false;
  std::string a = dir.empty() ? "fm_table.energy" : dir + "/fm_table.energy";
  std::string b = dir.empty() ? "fm_table.force" : dir + "/fm_table.force";
  FILE *fa = std::fopen(a.c_str(), "r"), *fb = std::fopen(b.c_str(), "r");
  bool ok = fa && fb;
  if (fa) std::fclose(fa);
  if (fb) std::fclose(fb);
  return ok;
}

// table mode, once the system is complete (fb.cpp:997-1017): the stored
// tables, or the tables built from the memories (written out when
// fm_use_table is set)
static AwStatus fm_table_setup(AwSystem &S, const std::string &dir)
{
  if (!S.P.fm_tb_flag) return AwStatus();
  if (fm_tables_stored(S.P, dir)) return aw_read_fm_table(dir, S);
  AwStatus s = aw_build_fm_table(S);
  if (s.ok() && S.P.fm_use_table) s = aw_write_fm_table(dir, S);
  return s;
}

static AwStatus load_system(const std::string &coeff, const std::string &seq_path, const std::string &dir, AwSystem &S,
                            bool from_files)
{
  std::string mem_file, gamma_file, cr_file, ff_file;
  AwStatus s = aw_parse_coeff(coeff, S.P, mem_file, gamma_file, cr_file, ff_file);
  if (!s.ok()) return s;
  if (from_files) {
    if (!(s = aw_read_sequence(seq_path, S)).ok()) return s;
  }
  if (!(s = aw_read_param_tables(dir, S)).ok()) return s;
  if (from_files) {
    if (!(s = aw_read_ssweight(dir, S)).ok()) return s;
  }
  if (S.P.fm_flag || S.P.fm_tb_flag) {
    std::string gp = gamma_file.empty() || gamma_file[0] == '/' ? gamma_file : dir + "/" + gamma_file;
    if (!(s = aw_read_gamma_array(gp, S.fmg)).ok()) return s;
    // table mode with stored tables reads no memories (fb.cpp:746-758)
    if (from_files && (S.P.fm_flag || !fm_tables_stored(S.P, dir)) && !(s = aw_read_mems(dir, mem_file, S)).ok())
      return s;
  }
  S.cr_file = cr_file;
  S.ff_file = ff_file;
  return AwStatus();
}

extern "C" {

int awsem_create(const char *coeff_path, const char *seq_path, const char *workdir, int precision, awsem_ctx **out)
{
  if (!out || !coeff_path || !seq_path) return AWSEM_E_ARG;
  *out = nullptr;
  std::unique_ptr<awsem_ctx> c(new awsem_ctx());
  c->precision = precision;
  std::string dir = workdir ? workdir : ".";
  c->dir = dir;
  AwStatus s = load_system(coeff_path, seq_path, dir, c->sys, true);
  if (s.ok()) s = aw_read_extras(dir, c->sys.cr_file, c->sys);
  if (s.ok()) s = aw_read_decoy_mems(dir, c->sys);
  if (s.ok()) s = aw_finish_system(c->sys);
  if (s.ok()) s = fm_table_setup(c->sys, dir);
  if (!s.ok()) {
    // keep the message reachable for the caller through a context-less path
    *out = c.release();
    (*out)->err = s.msg;
    (*out)->sys.n = -1;
    return s.code;
  }
  *out = c.release();
  return AWSEM_OK;
}

int awsem_create_arrays(const char *coeff_path, const char *param_dir, int nch, const int *chain_len, const char *seq,
                        const double *aps, int n_struct, const int *struct_len, const double *struct_xyz_nm,
                        const char *struct_seq, int n_mem, const int *mem_struct, const int *mem_tpos,
                        const int *mem_fpos, const int *mem_len, const double *mem_weight, int precision,
                        awsem_ctx **out)
{
  if (!out || !coeff_path || nch <= 0 || !chain_len || !seq) return AWSEM_E_ARG;
  *out = nullptr;
  std::unique_ptr<awsem_ctx> c(new awsem_ctx());
  c->precision = precision;
  AwSystem &S = c->sys;
  c->dir = param_dir ? param_dir : ".";
  AwStatus s = load_system(coeff_path, "", param_dir ? param_dir : ".", S, false);
  if (s.ok()) {
    S.nch = nch;
    int off = 0;
    for (int k = 0; k < nch; ++k) {
      if (chain_len[k] <= 0) { s = AwStatus::err(AWSEM_E_ARG, "chain length must be positive"); break; }
      S.ch_start.push_back(off);
      S.ch_len.push_back(chain_len[k]);
      off += chain_len[k];
    }
    S.n = off;
    if (s.ok()) {
      S.se.assign(seq, seq + S.n);
      for (char ch : S.se)
        if (ch < 'A' || ch > 'Z') { s = AwStatus::err(AWSEM_E_FORMAT, "Sequence: unknown residue letter"); break; }
    }
    if (s.ok()) {
      S.aps.assign((size_t)S.n * 12, 0.0);
      if (aps && S.P.ssweight_flag)
        for (int i = 0; i < S.n; ++i)
          for (int k = 0; k < 12; ++k)
            if (S.P.ssweight[k]) S.aps[(size_t)i * 12 + k] = aps[(size_t)i * 12 + k];
    }
    if (s.ok() && (S.P.fm_flag || S.P.fm_tb_flag)) {
      if (n_mem > 0 && (!struct_len || !struct_xyz_nm || !struct_seq || !mem_struct || !mem_tpos || !mem_fpos ||
                        !mem_len || !mem_weight))
        s = AwStatus::err(AWSEM_E_ARG, "memory arrays missing");
      std::vector<long long> soff(n_struct + 1, 0);
      for (int k = 0; k < n_struct && s.ok(); ++k) soff[k + 1] = soff[k] + struct_len[k];
      for (int k = 0; k < n_mem && s.ok(); ++k) {
        int st = mem_struct[k], tp = mem_tpos[k], fp = mem_fpos[k], len = mem_len[k];
        if (st < 0 || st >= n_struct || len <= 0 || fp < 0 || fp + len > struct_len[st])
          s = AwStatus::err(AWSEM_E_FORMAT, "read_mems: Fragment_Memory: Error reading memory fragment");
        else if (tp < 0 || tp + len > S.n)
          s = AwStatus::err(AWSEM_E_FORMAT, "read_mems: Fragment_Memory: Incorrectly defined memory fragment");
        if (!s.ok()) break;
        AwMemory m;
        m.pos = tp; m.len = len; m.weight = mem_weight[k];
        m.se.resize(len);
        m.ca.assign((size_t)len * 3, 0.0);
        m.cb.assign((size_t)len * 3, 0.0);
        for (int r = 0; r < len; ++r) {
          long long q = soff[st] + fp + r;
          m.se[r] = struct_seq[q];
          for (int l = 0; l < 3; ++l) {
            m.ca[3 * r + l] = struct_xyz_nm[(q * 2 + 0) * 3 + l] * 10;
            if (m.se[r] != 'G') m.cb[3 * r + l] = struct_xyz_nm[(q * 2 + 1) * 3 + l] * 10;
          }
        }
        S.mems.push_back(std::move(m));
      }
    }
  }
  if (s.ok()) s = aw_read_extras(param_dir ? param_dir : ".", S.cr_file, S);
  if (s.ok()) s = aw_read_decoy_mems(param_dir ? param_dir : ".", S);
  if (s.ok()) s = aw_finish_system(S);
  if (s.ok()) s = fm_table_setup(S, param_dir ? param_dir : ".");
  *out = c.release();
  if (!s.ok()) {
    (*out)->err = s.msg;
    (*out)->sys.n = -1;
    return s.code;
  }
  return AWSEM_OK;
}

int awsem_destroy(awsem_ctx *c)
{
  delete c;
  return AWSEM_OK;
}

int awsem_last_error(const awsem_ctx *c, char *buf, size_t n)
{
  if (!c || !buf || n == 0) return AWSEM_E_ARG;
  std::snprintf(buf, n, "%s", c->err.c_str());
  return AWSEM_OK;
}

int awsem_n_residues(const awsem_ctx *c, int *n)
{
  if (!c || !n) return AWSEM_E_ARG;
  *n = c->sys.n;
  return AWSEM_OK;
}

double awsem_pair_list_cutoff(const awsem_ctx *c) { return c ? c->sys.P.pair_list_cutoff : 0.0; }

int awsem_set_vexcluded(awsem_ctx *c, int enable, int p, const double *lambda, const double *cs, const double *cl)
{
  if (!c || c->sys.n < 0) return AWSEM_E_ARG;
  FFParams &P = c->sys.P;
  P.ev_on = enable ? 1 : 0;
  if (enable && lambda && cs && cl) {
    P.ev_p = p;
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b) {
        P.ev_lambda[a][b] = lambda[a * 6 + b];
        P.ev_cut_short[a][b] = cs[a * 6 + b];
        P.ev_cut_long[a][b] = cl[a * 6 + b];
      }
  }
  AwStatus s = aw_finish_system(c->sys);
  c->dirty = true;
  return s.ok() ? AWSEM_OK : fail(c, s);
}

int awsem_set_bonds(awsem_ctx *c, int enable, double k, const double *r0)
{
  if (!c || c->sys.n < 0) return AWSEM_E_ARG;
  FFParams &P = c->sys.P;
  P.bond_on = enable ? 1 : 0;
  if (enable) {
    P.bond_k = k;
    if (r0) for (int q = 0; q < 5; ++q) P.bond_r0[q] = r0[q];
  }
  c->dirty = true;
  return AWSEM_OK;
}

int awsem_set_list_skin(awsem_ctx *c, double skin)
{
  if (!c || c->sys.n < 0 || !(skin >= 0.0)) return AWSEM_E_ARG;
  c->sys.P.list_skin = skin;
  AwStatus s = aw_finish_system(c->sys);
  c->dirty = true;
  return s.ok() ? AWSEM_OK : fail(c, s);
}

int awsem_set_term_mask(awsem_ctx *c, uint32_t mask)
{
  if (!c || c->sys.n < 0) return AWSEM_E_ARG;
  c->sys.P.term_mask = mask;
  c->dirty = true;
  return AWSEM_OK;
}

// Construct_Computational_Arrays, fb.cpp:1590-1698, single rank (every atom
// local): residues are the distinct residue tags in increasing order; the
// library requires them to be 1..n, each with one atom in each group.
int awsem_set_topology(awsem_ctx *c, int n_atoms, const int *group, const int *mol, const int *res)
{
  if (!c || c->sys.n < 0 || n_atoms < 0 || (n_atoms > 0 && (!group || !mol || !res))) return AWSEM_E_ARG;
  const AwSystem &S = c->sys;
  int n = S.n;
  int cnt[3] = {0, 0, 0};
  for (int a = 0; a < n_atoms; ++a) {
    int g = group[a];
    if (g < -1 || g > 2) return fail(c, AwStatus::err(AWSEM_E_ARG, "group code must be -1, 0, 1 or 2"));
    if (g >= 0) {
      cnt[g]++;
      if (res[a] <= 0) return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "Molecular tag must be positive in fix backbone"));
    }
  }
  if (cnt[0] != cnt[1] || cnt[1] != cnt[2])
    return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "All groups must contain the same # of atoms in fix backbone"));
  if (cnt[0] != n) return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "Number of CA atoms differs from the sequence length"));
  std::vector<int> map(3 * (size_t)n, -1);
  for (int a = 0; a < n_atoms; ++a) {
    int g = group[a];
    if (g < 0) continue;
    int r = res[a];
    if (r > n) return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "Residue tag is out of range"));
    int i = r - 1;
    if (map[(size_t)g * n + i] != -1) return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "Two atoms of one group share a residue tag"));
    map[(size_t)g * n + i] = a;
  }
  for (int i = 0; i < n; ++i) {
    int ca = map[i], cb = map[(size_t)n + i], o = map[2 * (size_t)n + i];
    if (ca == -1) return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "Missing neighbor atoms in fix backbone (Code 004)"));
    if (cb == -1 || o == -1) return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "Missing neighbor atoms in fix backbone (Code 001)"));
    int chain = mol[ca];
    if (chain <= 0 || chain > S.nch) return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "Chain tag is out of range"));
    if (mol[cb] != chain || mol[o] != chain)
      return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "Atoms in a residue have different chain tag"));
    int c0 = S.ch_start[chain - 1];
    if (i < c0 || i > c0 + S.ch_len[chain - 1] - 1) return fail(c, AwStatus::err(AWSEM_E_TOPOLOGY, "Residue tag is out of range"));
  }
  c->atom_of = map;
  c->n_atoms = n_atoms;
  c->topo = true;
  return AWSEM_OK;
}

static AwStatus ensure_dev(awsem_ctx *c)
{
  if (c->dev && !c->dirty) return AwStatus();
  c->dev.reset();
  std::vector<const AwSystem *> v{&c->sys};
  AwStatus s = make_dev_batch(v, c->precision, false, c->dev);
  if (s.ok()) c->dirty = false;
  return s;
}

static bool finite_all(const double *x, size_t n)
{
  for (size_t k = 0; k < n; ++k) if (!std::isfinite(x[k])) return false;
  return true;
}

}  // extern "C"

AwStatus aw_ctx_device(awsem_ctx *c, const double *xres, AwCtxDev &out)
{
  AwStatus s = ensure_dev(c);
  if (!s.ok()) return s;
  DevBatch &D = *c->dev;
  HIPCHK(hipMemcpyAsync(D.d_x, xres, (size_t)c->sys.n * 9 * sizeof(double), hipMemcpyHostToDevice, D.stream));
  // the context's coordinates changed: its list and forces are stale
  D.nl_dirty = true;
  D.f_valid = false;
  out.params = (const FFParams *)D.d_params;
  out.rinfo = D.reps[0].rinfo;
  out.x = D.d_x;
  out.stream = (void *)D.stream;
  return AwStatus();
}

AwFrustState &aw_ctx_frust(awsem_ctx *c) { return c->tf; }
const std::string &aw_ctx_dir(const awsem_ctx *c) { return c->dir; }
int aw_ctx_fail(awsem_ctx *c, const AwStatus &s) { return fail(c, s); }

extern "C" {

static AwStatus compute_res(awsem_ctx *c, const double *xres, double *fres, double *energy)
{
  AwStatus s = ensure_dev(c);
  if (!s.ok()) return s;
  DevBatch &D = *c->dev;
  size_t bytes = (size_t)c->sys.n * 9 * sizeof(double);
  HIPCHK(hipMemcpyAsync(D.d_x, xres, bytes, hipMemcpyHostToDevice, D.stream));
  AwRun run;
  std::memset(&run, 0, sizeof run);
  run.mode = 0;
  AwStatus ss = D.set_schedule(0, nullptr);
  if (!ss.ok()) return ss;
  HIPCHK(hipEventRecord(D.ev0, D.stream));
  if (aw_launch_phases(D.ptrs(), run, D.stream)) return AwStatus::err(AWSEM_E_DEVICE, "kernel launch failed");
  HIPCHK(hipEventRecord(D.ev1, D.stream));
  std::vector<double> f(c->sys.n * 9);
  HIPCHK(hipMemcpyAsync(f.data(), D.d_f, bytes, hipMemcpyDeviceToHost, D.stream));
  double e[AW_N_ET];
  HIPCHK(hipMemcpyAsync(e, D.d_e, sizeof e, hipMemcpyDeviceToHost, D.stream));
  HIPCHK(hipStreamSynchronize(D.stream));
  if (!(s = take_faults(D)).ok()) return s;
  float ms = 0.f;
  hipEventElapsedTime(&ms, D.ev0, D.ev1);
  D.last_ms = ms;
  for (size_t k = 0; k < f.size(); ++k) fres[k] += f[k];
  if (energy) for (int k = 0; k < AW_N_ET; ++k) energy[k] = e[k];
  for (int k = 0; k < AW_N_ET; ++k) c->last_e[k] = e[k];
  if (!std::isfinite(e[0])) return AwStatus::err(AWSEM_E_NONFINITE, "non-finite energy");
  return AwStatus();
}

int awsem_compute_res(awsem_ctx *c, const double *xres, double *fres, double *energy)
{
  if (!c || c->sys.n < 0 || !xres || !fres) return AWSEM_E_ARG;
  if (!finite_all(xres, (size_t)c->sys.n * 9)) return fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates"));
  AwStatus s = compute_res(c, xres, fres, energy);
  return s.ok() ? AWSEM_OK : fail(c, s);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the residue-layout evaluation over a caller-given device region (used by
// awsem_compute_res_noalloc, awsem_noalloc.hip): the region holds the
// coordinates, the forces, the energies and the phase engine's working slab,
// laid out as [x 9n][f 9n][energy 32][slab, 256-byte aligned]

static size_t noalloc_slab_off(int n) { return ((18 * (size_t)n + 32 + 31) / 32) * 32; }

size_t aw_noalloc_doubles(int n) { return noalloc_slab_off(n) + (size_t)aw_slab_doubles(n); }

int aw_noalloc_check(awsem_ctx *c, const double *xres, const double *fres, size_t cap_doubles)
{
  if (!c || c->sys.n < 0 || !xres || !fres) return AWSEM_E_ARG;
  int n = c->sys.n;
  if (aw_noalloc_doubles(n) > cap_doubles) {
    int nmax = 0;
    while (aw_noalloc_doubles(nmax + 1) <= cap_doubles) ++nmax;
    char msg[256];
    std::snprintf(msg, sizeof msg,
                  "awsem_compute_res_noalloc: %d residues need %zu bytes of scratch; its fixed scratch array holds "
                  "%zu bytes (at most %d residues)", n, aw_noalloc_doubles(n) * sizeof(double),
                  cap_doubles * sizeof(double), nmax);
    return fail(c, AwStatus::err(AWSEM_E_NOMEM, msg));
  }
  if (!c->dev || c->dirty)
    return fail(c, AwStatus::err(AWSEM_E_STATE, "awsem_compute_res_noalloc: the context's device image is not built "
                                                "(or its parameters changed): call awsem_compute_res, awsem_compute "
                                                "or awsem_run once first"));
  if (!finite_all(xres, (size_t)n * 9)) return fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates"));
  return AWSEM_OK;
}

int aw_compute_res_scratch(awsem_ctx *c, const double *xres, double *fres, double *energy, double *dscr)
{
  DevBatch &D = *c->dev;
  int n = c->sys.n;
  size_t n9 = (size_t)n * 9;
  // the context's descriptor with its per-call buffers moved into the region
  ReplicaDesc R = D.reps[0];
  R.x = dscr;
  R.f = dscr + n9;
  R.energy = dscr + 2 * n9;
  R.scratch = dscr + noalloc_slab_off(n);
  R.nsteps = 0; R.ramp0 = 0; R.run_steps = 0; R.rng_id = 0; R.flags = 0;
  AwRun run;
  std::memset(&run, 0, sizeof run);
  run.mode = 0;
  auto dev_fail = [&](const char *what) {
    (void)hipStreamSynchronize(D.stream);
    // the context's own descriptor back in place for its other entry points
    (void)hipMemcpy(D.d_rep, D.reps.data(), sizeof(ReplicaDesc), hipMemcpyHostToDevice);
    return fail(c, AwStatus::err(AWSEM_E_DEVICE, what));
  };
  if (hipMemcpyAsync(R.x, xres, n9 * sizeof(double), hipMemcpyHostToDevice, D.stream) != hipSuccess ||
      hipMemcpyAsync(D.d_rep, &R, sizeof R, hipMemcpyHostToDevice, D.stream) != hipSuccess)
    return dev_fail("awsem_compute_res_noalloc: upload");
  if (aw_launch_phases(D.ptrs(), run, D.stream)) return dev_fail("awsem_compute_res_noalloc: kernel launch failed");
  // every later call of the context's other entry points re-reads its own
  // descriptor (awsem_list reads it without rewriting it first)
  if (hipMemcpyAsync(D.d_rep, D.reps.data(), sizeof(ReplicaDesc), hipMemcpyHostToDevice, D.stream) != hipSuccess)
    return dev_fail("awsem_compute_res_noalloc: descriptor restore");
  double e[AW_N_ET];
  if (hipMemcpyAsync(e, R.energy, sizeof e, hipMemcpyDeviceToHost, D.stream) != hipSuccess)
    return dev_fail("awsem_compute_res_noalloc: energy copy");
  // forces added into fres through a fixed host buffer (no allocation)
  double buf[2048];
  for (size_t k0 = 0; k0 < n9; k0 += 2048) {
    size_t m = std::min<size_t>(2048, n9 - k0);
    if (hipMemcpyAsync(buf, R.f + k0, m * sizeof(double), hipMemcpyDeviceToHost, D.stream) != hipSuccess ||
        hipStreamSynchronize(D.stream) != hipSuccess)
      return dev_fail("awsem_compute_res_noalloc: force copy");
    for (size_t k = 0; k < m; ++k) fres[k0 + k] += buf[k];
  }
  if (hipStreamSynchronize(D.stream) != hipSuccess) return dev_fail("awsem_compute_res_noalloc: sync");
  {
    AwStatus fs = take_faults(D);
    if (!fs.ok()) return fail(c, fs);
  }
  if (energy) for (int k = 0; k < AW_N_ET; ++k) energy[k] = e[k];
  for (int k = 0; k < AW_N_ET; ++k) c->last_e[k] = e[k];
  if (!std::isfinite(e[0])) return fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite energy"));
  return AWSEM_OK;
}

extern "C" {

// the atom-layout evaluation: gather CA/CB/O of every residue from the
// caller's atom order (unwrapped with the image flags in a periodic box,
// fb.cpp:7067-7112), evaluate, scatter-add the forces back
static int compute_atoms(awsem_ctx *c, const double *x, const int32_t *image, double *f, double *energy)
{
  if (!c->topo) return fail(c, AwStatus::err(AWSEM_E_STATE, "awsem_set_topology must be called first"));
  int n = c->sys.n;
  std::vector<double> xr((size_t)n * 9), fr((size_t)n * 9, 0.0);
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < n; ++i) {
      int a = c->atom_of[(size_t)k * n + i];
      double *dst = &xr[((size_t)k * n + i) * 3];
      for (int l = 0; l < 3; ++l) dst[l] = x[(size_t)a * 3 + l];
      if (image) {
        int img = image[a];
        int box[3] = {(img & 1023) - 512, (img >> 10 & 1023) - 512, (img >> 20) - 512};
        for (int l = 0; l < 3; ++l)
          if (c->periodic[l]) dst[l] = x[(size_t)a * 3 + l] + box[l] * c->prd[l];
      }
    }
  if (!finite_all(xr.data(), xr.size())) return fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates"));
  AwStatus s = compute_res(c, xr.data(), fr.data(), energy);
  if (!s.ok()) return fail(c, s);
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < n; ++i) {
      int a = c->atom_of[(size_t)k * n + i];
      for (int l = 0; l < 3; ++l) f[(size_t)a * 3 + l] += fr[((size_t)k * n + i) * 3 + l];
    }
  return AWSEM_OK;
}

int awsem_compute(awsem_ctx *c, const double *x, double *f, double *energy)
{
  if (!c || c->sys.n < 0 || !x || !f) return AWSEM_E_ARG;
  if (c->periodic[0] || c->periodic[1] || c->periodic[2])
    return fail(c, AwStatus::err(AWSEM_E_UNSUPPORTED, "a periodic box needs the atoms' image flags: use awsem_compute_image"));
  return compute_atoms(c, x, nullptr, f, energy);
}

int awsem_compute_image(awsem_ctx *c, const double *x, const int32_t *image, double *f, double *energy)
{
  if (!c || c->sys.n < 0 || !x || !f || !image) return AWSEM_E_ARG;
  return compute_atoms(c, x, image, f, energy);
}

int awsem_set_box(awsem_ctx *c, const int *periodic, const double *prd)
{
  if (!c || c->sys.n < 0 || !periodic) return AWSEM_E_ARG;
  for (int l = 0; l < 3; ++l)
    if (periodic[l] && (!prd || !(prd[l] > 0.0) || !std::isfinite(prd[l])))
      return fail(c, AwStatus::err(AWSEM_E_ARG, "a periodic dimension needs a positive box length"));
  for (int l = 0; l < 3; ++l) {
    c->periodic[l] = periodic[l] ? 1 : 0;
    c->prd[l] = periodic[l] ? prd[l] : 0.0;
  }
  return AWSEM_OK;
}

// energy.log, fb.cpp:101-106 (header) and 7661-7672 (one line per thermo
// step): the step, then "\t%8.6f" for the seventeen FixBackbone terms in
// EnergyTerms order and the total.  The fix's own excluded-volume slot is
// zero in the reference's production build (its terms run under
// DEBUGFORCES only, fb.cpp:7396-7403); the excluded-volume pair style and
// the bonds are LAMMPS styles that report their energies in LAMMPS' thermo
// output, so this log leaves them out of the columns and of the total.
int awsem_energy_log_open(awsem_ctx *c, const char *path)
{
  if (!c || c->sys.n < 0 || !path) return AWSEM_E_ARG;
  if (c->elog) std::fclose(c->elog);
  c->elog = std::fopen(path, "w");
  if (!c->elog) return fail(c, AwStatus::err(AWSEM_E_FILE, std::string("cannot open ") + path));
  std::fputs("Step   \tChain   \tShake   \tChi     \tRama    \tExcluded\tDSSP    \tP_AP    \tWater   \tBurial  "
             "\tHelix   \tAMH-Go  \tFrag_Mem\tVec_FM  \tContact_Restraints  \tMembrane\tSSB     \tElectro.\tVTotal\n",
             c->elog);
  std::fflush(c->elog);
  return AWSEM_OK;
}

int awsem_energy_log_write(awsem_ctx *c, int64_t step, const double *energy)
{
  if (!c || c->sys.n < 0) return AWSEM_E_ARG;
  if (!c->elog) return fail(c, AwStatus::err(AWSEM_E_STATE, "awsem_energy_log_open must be called first"));
  const double *e = energy ? energy : c->last_e;
  double col[18], total = 0.0;
  for (int k = 1; k < 18; ++k) {
    col[k] = k == AW_ET_VEXCLUDED ? 0.0 : e[k];
    total += col[k];
  }
  std::fprintf(c->elog, "%lld ", (long long)step);
  for (int k = 1; k < 18; ++k) std::fprintf(c->elog, "\t%8.6f", col[k]);
  std::fprintf(c->elog, "\t%8.6f\n", total);
  std::fflush(c->elog);
  return AWSEM_OK;
}

int awsem_energy_log_close(awsem_ctx *c)
{
  if (!c) return AWSEM_E_ARG;
  if (c->elog) std::fclose(c->elog);
  c->elog = nullptr;
  return AWSEM_OK;
}

int awsem_list(awsem_ctx *c, int kind, int arg, const double *xres, int64_t *out, int cap, int *count)
{
  if (!c || c->sys.n < 0 || !xres || !count || kind < 0 || kind > 6 || cap < 0) return AWSEM_E_ARG;
  AwStatus s = ensure_dev(c);
  if (!s.ok()) return fail(c, s);
  DevBatch &D = *c->dev;
  if (D.list_cap < std::max(cap, 1)) {
    if (D.d_list) hipFree(D.d_list);
    D.list_cap = std::max(cap, 1);
    if (hipMalloc((void **)&D.d_list, (size_t)D.list_cap * sizeof(long long)) != hipSuccess) return fail(c, AwStatus::err(AWSEM_E_NOMEM, "list buffer"));
  }
  size_t bytes = (size_t)c->sys.n * 9 * sizeof(double);
  if (hipMemcpyAsync(D.d_x, xres, bytes, hipMemcpyHostToDevice, D.stream) != hipSuccess) return fail(c, AwStatus::err(AWSEM_E_DEVICE, "copy"));
  if (aw_launch_list(D.ptrs(), 0, kind, arg, D.d_list, D.list_cap, D.d_count, D.stream)) return fail(c, AwStatus::err(AWSEM_E_DEVICE, "list launch"));
  int cnt = 0;
  if (hipMemcpyAsync(&cnt, D.d_count, sizeof(int), hipMemcpyDeviceToHost, D.stream) != hipSuccess ||
      hipStreamSynchronize(D.stream) != hipSuccess)
    return fail(c, AwStatus::err(AWSEM_E_DEVICE, "list copy"));
  *count = cnt;
  if (cnt > cap) return fail(c, AwStatus::err(AWSEM_E_NOMEM, "list capacity too small"));
  std::vector<long long> tmp(cnt);
  if (cnt && hipMemcpy(tmp.data(), D.d_list, (size_t)cnt * sizeof(long long), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(c, AwStatus::err(AWSEM_E_DEVICE, "list copy"));
  std::sort(tmp.begin(), tmp.end());
  for (int k = 0; k < cnt; ++k) out[k] = tmp[k];
  return AWSEM_OK;
}

int awsem_run(awsem_ctx *c, int nsteps, const awsem_langevin *lg, double *xres, double *vres, double *fres, double *energy)
{
  if (!c || c->sys.n < 0 || !lg || !xres || !vres || nsteps < 0) return AWSEM_E_ARG;
  AwStatus s = ensure_dev(c);
  if (!s.ok()) return fail(c, s);
  DevBatch &D = *c->dev;
  size_t bytes = (size_t)c->sys.n * 9 * sizeof(double);
  if (hipMemcpyAsync(D.d_x, xres, bytes, hipMemcpyHostToDevice, D.stream) != hipSuccess ||
      hipMemcpyAsync(D.d_v, vres, bytes, hipMemcpyHostToDevice, D.stream) != hipSuccess)
    return fail(c, AwStatus::err(AWSEM_E_DEVICE, "copy"));
  AwRun run;
  std::memset(&run, 0, sizeof run);
  run.mode = 1; run.nsteps = nsteps; run.dt = lg->dt; run.damp = lg->damp;
  // the coordinates come from the caller: the Verlet list is rebuilt and
  // the first forces evaluated before the first step
  D.nl_dirty = true;
  D.f_valid = false;
  s = D.set_schedule(nsteps, lg);
  if (!s.ok()) return fail(c, s);
  (void)hipEventRecord(D.ev0, D.stream);
  // the step sequence replayed from a captured graph, on the phase engine
  // (all-pairs sweeps, eight launches per step) or, with
  // AWSEM_RUN_ENGINE=wide, the wide engine (Verlet list and work queues,
  // eleven launches per step).  For one system the phase engine is the
  // faster: 12.5k / 7.4k / 2.9k steps/s at 200 / 1000 / 3000 residues
  // against 9.2k / 5.4k / 2.9k (FP64, MI355X)
  int eng = 2;
  if (const char *e = std::getenv("AWSEM_RUN_ENGINE")) eng = std::strcmp(e, "wide") == 0 ? 1 : 2;
  if (!(s = launch_seq(D, run, eng, nsteps)).ok()) return fail(c, s);
  (void)hipEventRecord(D.ev1, D.stream);
  double e[AW_N_ET];
  if (hipMemcpyAsync(xres, D.d_x, bytes, hipMemcpyDeviceToHost, D.stream) != hipSuccess ||
      hipMemcpyAsync(vres, D.d_v, bytes, hipMemcpyDeviceToHost, D.stream) != hipSuccess ||
      (fres && hipMemcpyAsync(fres, D.d_f, bytes, hipMemcpyDeviceToHost, D.stream) != hipSuccess) ||
      hipMemcpyAsync(e, D.d_e, sizeof e, hipMemcpyDeviceToHost, D.stream) != hipSuccess ||
      hipStreamSynchronize(D.stream) != hipSuccess)
    return fail(c, AwStatus::err(AWSEM_E_DEVICE, "copy back"));
  if (!(s = take_faults(D)).ok()) return fail(c, s);
  float ms = 0.f;
  hipEventElapsedTime(&ms, D.ev0, D.ev1);
  D.last_ms = ms;
  if (energy) for (int k = 0; k < AW_N_ET; ++k) energy[k] = e[k];
  for (int k = 0; k < AW_N_ET; ++k) c->last_e[k] = e[k];
  if (!std::isfinite(e[0])) return fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite energy"));
  return AWSEM_OK;
}

// ---------------------------------------------------------------------------
// batches

static int bfail(awsem_batch *b, const AwStatus &s)
{
  if (b) b->err = s.msg;
  return s.code;
}

int awsem_batch_create(int n_ctx, awsem_ctx *const *ctxs, int precision, awsem_batch **out)
{
  if (!out || n_ctx <= 0 || !ctxs) return AWSEM_E_ARG;
  *out = nullptr;
  std::vector<const AwSystem *> v;
  for (int k = 0; k < n_ctx; ++k) {
    if (!ctxs[k] || ctxs[k]->sys.n < 0) return AWSEM_E_ARG;
    v.push_back(&ctxs[k]->sys);
  }
  std::unique_ptr<awsem_batch> b(new awsem_batch());
  for (const AwSystem *S : v)
    if (S->n >= 32768) {
      // residue pairs are packed as (i << 16) | j in the batch engine's lists
      b->err = "batched systems are limited to 32767 residues each";
      *out = b.release();
      return AWSEM_E_ARG;
    }
  AwStatus s = make_dev_batch(v, precision, true, b->dev);
  if (!s.ok()) { *out = b.release(); (*out)->err = s.msg; return s.code; }
  *out = b.release();
  return AWSEM_OK;
}

int awsem_batch_size(const awsem_batch *b, int *nrep, int *nres)
{
  if (!b || !b->dev) return AWSEM_E_ARG;
  if (nrep) *nrep = b->dev->nrep;
  if (nres) *nres = b->dev->total_res;
  return AWSEM_OK;
}

int awsem_batch_set_state(awsem_batch *b, const double *x, const double *v)
{
  if (!b || !b->dev) return AWSEM_E_ARG;
  DevBatch &D = *b->dev;
  size_t bytes = (size_t)D.total_res * 9 * sizeof(double);
  if ((x && !finite_all(x, (size_t)D.total_res * 9)) || (v && !finite_all(v, (size_t)D.total_res * 9)))
    return bfail(b, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates or velocities"));
  if (x && hipMemcpy(D.d_x, x, bytes, hipMemcpyHostToDevice) != hipSuccess) return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "copy x"));
  if (x) { D.nl_dirty = true; D.f_valid = false; }
  if (v && hipMemcpy(D.d_v, v, bytes, hipMemcpyHostToDevice) != hipSuccess) return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "copy v"));
  return AWSEM_OK;
}

// compute qwolynes / qonuchic of replicas [first, first + count) at their
// current positions on the device (no copy of the state)
int awsem_batch_q(awsem_batch *b, awsem_q *q, int first, int count, double *values)
{
  if (!b || !b->dev || !q || !values || first < 0 || count <= 0 || first + count > b->dev->nrep) return AWSEM_E_ARG;
  DevBatch &D = *b->dev;
  std::vector<long long> off(count);
  for (int k = 0; k < count; ++k) {
    int r = first + k;
    if (D.reps[r].n != aw_q_residues(q))
      return bfail(b, AwStatus::err(AWSEM_E_ARG, "order parameter: replica residue count differs from its native"));
    off[k] = 9LL * D.res_off[r];
  }
  int rc = aw_q_eval(q, D.d_x, off.data(), count, values, D.stream);
  if (rc) return bfail(b, AwStatus::err(rc, "order parameter evaluation failed"));
  return AWSEM_OK;
}

int awsem_batch_get_state(awsem_batch *b, double *x, double *v, double *f, double *energy)
{
  if (!b || !b->dev) return AWSEM_E_ARG;
  DevBatch &D = *b->dev;
  if (hipStreamSynchronize(D.stream) != hipSuccess) return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "sync"));
  size_t bytes = (size_t)D.total_res * 9 * sizeof(double);
  if ((x && hipMemcpy(x, D.d_x, bytes, hipMemcpyDeviceToHost) != hipSuccess) ||
      (v && hipMemcpy(v, D.d_v, bytes, hipMemcpyDeviceToHost) != hipSuccess) ||
      (f && hipMemcpy(f, D.d_f, bytes, hipMemcpyDeviceToHost) != hipSuccess) ||
      (energy && hipMemcpy(energy, D.d_e, (size_t)D.nrep * AW_N_ET * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess))
    return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "copy back"));
  return AWSEM_OK;
}

// one launch sequence of a device batch on its stream: replayed from a
// captured HIP graph (captured again when the mode, dt, damp, engine or,
// for the wide and phase engines, the step count change), or launched
// directly
static AwStatus launch_seq(DevBatch &D, const AwRun &run, int engine, int nst)
{
  auto launch = [&]() {
    if (engine == 2) return aw_launch_phases(D.ptrs(), run, D.stream);
    return engine == 1 ? aw_launch_wide(D.ptrs(), run, nst, D.stream) : aw_launch_resident(D.ptrs(), run, D.stream);
  };
  if (!D.use_graph) {
    if (launch()) return AwStatus::err(AWSEM_E_DEVICE, "kernel launch failed");
    return AwStatus();
  }
  // the wide and phase engines' launch sequences depend on the step count
  int gkey_steps = engine >= 1 ? nst : -1;
  if (!D.graph || D.g_mode != run.mode || D.g_dt != run.dt || D.g_damp != run.damp || D.g_engine != engine ||
      D.g_nsteps != gkey_steps) {
    if (D.graph) { (void)hipGraphExecDestroy(D.graph); D.graph = nullptr; }
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(D.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return AwStatus::err(AWSEM_E_DEVICE, "graph capture");
    int lrc = launch();
    hipError_t ce = hipStreamEndCapture(D.stream, &g);
    if (lrc || ce != hipSuccess || !g) return AwStatus::err(AWSEM_E_DEVICE, "graph capture of the launch sequence");
    hipError_t ie = hipGraphInstantiate(&D.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) { D.graph = nullptr; return AwStatus::err(AWSEM_E_DEVICE, "graph instantiate"); }
    D.g_mode = run.mode; D.g_dt = run.dt; D.g_damp = run.damp; D.g_engine = engine; D.g_nsteps = gkey_steps;
  }
  if (hipGraphLaunch(D.graph, D.stream) != hipSuccess) return AwStatus::err(AWSEM_E_DEVICE, "graph launch");
  return AwStatus();
}

static int batch_launch(awsem_batch *b, AwRun run, const awsem_langevin *lg)
{
  DevBatch &D = *b->dev;
  AwStatus s = D.set_schedule(run.mode == 1 ? run.nsteps : 0, lg);
  if (!s.ok()) return bfail(b, s);
  D.nl_dirty = false;
  D.f_valid = false;
  (void)hipEventRecord(D.ev0, D.stream);
  int nst = run.mode == 1 ? run.nsteps : 0;
  if (!(s = launch_seq(D, run, D.engine, nst)).ok()) return bfail(b, s);
  (void)hipEventRecord(D.ev1, D.stream);
  hipError_t e = hipEventSynchronize(D.ev1);
  if (e != hipSuccess) return bfail(b, AwStatus::err(AWSEM_E_DEVICE, hipGetErrorString(e)));
  if (!(s = take_faults(D)).ok()) return bfail(b, s);
  if (D.nshare > 0 && D.engine == 0) {
    // a replica that gave up waiting for chunks taken elsewhere (bounded
    // spin) leaves this flag: the launch's forces are then incomplete
    int flag = 0;
    if (hipMemcpy(&flag, D.d_queue + 32, sizeof flag, hipMemcpyDeviceToHost) != hipSuccess)
      return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "reading the sharing flag"));
    if (flag) {
      // claim words, counters and accumulators back to their initial state
      (void)hipMemset(D.d_queue + 32, 0, sizeof flag);
      (void)hipMemcpy(D.d_share, D.share_img.data(), D.share_img.size(), hipMemcpyHostToDevice);
      return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "fragment-memory sharing: a replica gave up waiting for its "
                                                    "chunks; the launch's results are invalid"));
    }
  }
  D.f_valid = true;
  D.last_engine = D.engine;
  float ms = 0.f;
  hipEventElapsedTime(&ms, D.ev0, D.ev1);
  b->ms_total = ms;
  b->n_kernels = 1;
  return AWSEM_OK;
}

int awsem_batch_compute(awsem_batch *b)
{
  if (!b || !b->dev) return AWSEM_E_ARG;
  AwRun run;
  std::memset(&run, 0, sizeof run);
  run.mode = 0;
  return batch_launch(b, run, nullptr);
}

int awsem_batch_run(awsem_batch *b, int nsteps, const awsem_langevin *lg)
{
  if (!b || !b->dev || !lg || nsteps < 0) return AWSEM_E_ARG;
  AwRun run;
  std::memset(&run, 0, sizeof run);
  run.mode = 1; run.nsteps = nsteps; run.dt = lg->dt; run.damp = lg->damp;
  return batch_launch(b, run, lg);
}

int awsem_batch_use_graph(awsem_batch *b, int on)
{
  if (!b || !b->dev) return AWSEM_E_ARG;
  b->dev->use_graph = on != 0;
  return AWSEM_OK;
}

int awsem_batch_set_engine(awsem_batch *b, int engine)
{
  if (!b || !b->dev || engine < 0 || engine > 1) return AWSEM_E_ARG;
  b->dev->engine = engine;
  return AWSEM_OK;
}

int awsem_batch_set_skin(awsem_batch *b, double skin)
{
  if (!b || !b->dev || !(skin > 0.0) || skin > 20.0) return AWSEM_E_ARG;
  DevBatch &D = *b->dev;
  // a list built with a larger skin stays a superset of every selectable
  // pair until some residue drifts by the new (smaller) skin/2, which is
  // when the drift test rebuilds it; only a larger skin needs a rebuild now
  if (skin > D.P.nl_skin) D.nl_dirty = true;
  D.P.nl_skin = skin;
  if (hipMemcpy(D.d_params, &D.P, sizeof(FFParams), hipMemcpyHostToDevice) != hipSuccess)
    return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "parameter upload"));
  return AWSEM_OK;
}

// Times short dynamics runs of this batch for each (skin, steps per launch)
// candidate, from the current state each time, and keeps the fastest skin.
// The state (x, v, f) is restored afterwards.
int awsem_batch_autotune(awsem_batch *b, const awsem_langevin *lg, int probe_steps, double *skin_out,
                         int *steps_per_launch_out, int *engine_out, double *ms_per_step_out)
{
  if (!b || !b->dev || !lg || probe_steps <= 0) return AWSEM_E_ARG;
  DevBatch &D = *b->dev;
  size_t bytes = (size_t)D.total_res * 9 * sizeof(double);
  double *save = nullptr;
  if (hipMalloc((void **)&save, 3 * std::max<size_t>(bytes, 8)) != hipSuccess) return bfail(b, AwStatus::err(AWSEM_E_NOMEM, "autotune"));
  auto copy3 = [&](bool to_save) {
    double *src[3] = {D.d_x, D.d_v, D.d_f};
    for (int k = 0; k < 3; ++k) {
      double *a = save + k * (bytes / sizeof(double)), *c = src[k];
      (void)hipMemcpyAsync(to_save ? a : c, to_save ? c : a, bytes, hipMemcpyDeviceToDevice, D.stream);
    }
    (void)hipStreamSynchronize(D.stream);
  };
  copy3(true);
  bool fv0 = D.f_valid;
  double skin0 = D.P.nl_skin;
  const double skins[] = {1.0, 1.5, 2.0, 2.5, 3.0};
  const int chunks[] = {5, 10, 20, 40, 80, 160};
  double best = 1e300, best_skin = skin0;
  int best_chunk = 10, best_engine = D.engine;
  int rc = AWSEM_OK;
  for (int eng = 0; eng < 2 && rc == AWSEM_OK; ++eng)
  for (double sk : skins) {
    D.engine = eng;
    if ((rc = awsem_batch_set_skin(b, sk)) != AWSEM_OK) break;
    for (int ch : chunks) {
      copy3(false);
      D.f_valid = fv0;
      D.nl_dirty = true;
      awsem_langevin l = *lg;
      if ((rc = awsem_batch_run(b, ch, &l)) != AWSEM_OK) break;  // warm-up: lists, caches
      double ms = 0.0;
      int done = 0;
      while (done < probe_steps) {
        l.step0 += ch;
        if ((rc = awsem_batch_run(b, ch, &l)) != AWSEM_OK) break;
        ms += b->ms_total;
        done += ch;
      }
      if (rc != AWSEM_OK) break;
      double per = ms / done;
      if (per < best) { best = per; best_skin = sk; best_chunk = ch; best_engine = eng; }
    }
    if (rc != AWSEM_OK) break;
  }
  copy3(false);
  (void)hipFree(save);
  D.f_valid = fv0;
  D.engine = best_engine;
  int rc2 = awsem_batch_set_skin(b, rc == AWSEM_OK ? best_skin : skin0);
  if (rc != AWSEM_OK) return rc;
  if (rc2 != AWSEM_OK) return rc2;
  if (skin_out) *skin_out = best_skin;
  if (steps_per_launch_out) *steps_per_launch_out = best_chunk;
  if (engine_out) *engine_out = best_engine;
  if (ms_per_step_out) *ms_per_step_out = best;
  return AWSEM_OK;
}

int awsem_batch_timing(const awsem_batch *b, double *ms_total, double *ms_kernel_avg, int *n_kernels)
{
  if (!b) return AWSEM_E_ARG;
  if (ms_total) *ms_total = b->ms_total;
  if (ms_kernel_avg) *ms_kernel_avg = b->n_kernels ? b->ms_total / b->n_kernels : 0.0;
  if (n_kernels) *n_kernels = b->n_kernels;
  return AWSEM_OK;
}

int awsem_batch_phase_clock(const awsem_batch *b, uint64_t *ticks32)
{
  if (!b || !b->dev || !ticks32) return AWSEM_E_ARG;
  if (!b->dev->d_prof) return AWSEM_E_STATE;
  if (hipMemcpy(ticks32, b->dev->d_prof, 32 * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess) return AWSEM_E_DEVICE;
  return AWSEM_OK;
}

int awsem_batch_list_stats(const awsem_batch *b, long long *builds, long long *pairs, int *n_invalid)
{
  if (!b || !b->dev) return AWSEM_E_ARG;
  const DevBatch &D = *b->dev;
  if (hipStreamSynchronize(D.stream) != hipSuccess) return AWSEM_E_DEVICE;
  long long nb = 0, np = 0;
  int bad = 0;
  for (const ReplicaDesc &R : D.reps) {
    int h[4] = {0, 0, 0, 0};
    if (R.nl_cap > 0 && hipMemcpy(h, R.nl, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
      return AWSEM_E_DEVICE;
    nb += h[2];
    np += h[0];
    bad += h[1] ? 0 : 1;
  }
  if (builds) *builds = nb;
  if (pairs) *pairs = np;
  if (n_invalid) *n_invalid = bad;
  return AWSEM_OK;
}

int awsem_batch_share_stats(const awsem_batch *b, int *n_shared, long long *chunks)
{
  if (!b || !b->dev) return AWSEM_E_ARG;
  const DevBatch &D = *b->dev;
  if (hipStreamSynchronize(D.stream) != hipSuccess) return AWSEM_E_DEVICE;
  int c = 0;
  if (D.nshare > 0 && hipMemcpy(&c, D.d_queue + 48, sizeof c, hipMemcpyDeviceToHost) != hipSuccess)
    return AWSEM_E_DEVICE;
  if (n_shared) *n_shared = D.nshare;
  if (chunks) *chunks = c;
  return AWSEM_OK;
}

int awsem_batch_lists(awsem_batch *b, int replica, int kind, int64_t *out, int cap, int *count)
{
  if (!b || !b->dev || !count || kind < 0 || kind > 4 || cap < 0 || (cap > 0 && !out)) return AWSEM_E_ARG;
  DevBatch &D = *b->dev;
  if (replica < 0 || replica >= D.nrep) return AWSEM_E_ARG;
  if (D.last_engine < 0) return bfail(b, AwStatus::err(AWSEM_E_STATE, "no evaluation has run on this batch"));
  if (hipStreamSynchronize(D.stream) != hipSuccess) return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "sync"));
  const ReplicaDesc &R = D.reps[replica];
  int h[AW_NL_HDR];
  if (R.nl_cap <= 0 || hipMemcpy(h, R.nl, sizeof h, hipMemcpyDeviceToHost) != hipSuccess)
    return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "list header copy"));
  if (!h[1]) return bfail(b, AwStatus::err(AWSEM_E_STATE, "the replica's Verlet list is not valid (overflowed or never built)"));
  // queue lengths: the wide engine keeps its counters in header slots 4-11
  // (queues at 6-9), the resident engine stores them to slots 12-15
  int len;
  if (kind == 0) len = h[0];
  else len = D.last_engine == 1 ? h[5 + kind] : h[11 + kind];
  if (len < 0 || len > (kind == 1 ? 2 * R.nl_cap : R.nl_cap)) return bfail(b, AwStatus::err(AWSEM_E_STATE, "queue length out of range"));
  *count = len;
  if (len > cap) return bfail(b, AwStatus::err(AWSEM_E_NOMEM, "list capacity too small"));
  const int qoff[5] = {0, R.nl_cap, 3 * R.nl_cap, 4 * R.nl_cap, 5 * R.nl_cap};
  std::vector<int> packed(std::max(len, 1));
  if (len && hipMemcpy(packed.data(), R.nl + AW_NL_HDR + qoff[kind], (size_t)len * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return bfail(b, AwStatus::err(AWSEM_E_DEVICE, "list copy"));
  long long n = R.n;
  std::vector<long long> codes(len);
  for (int k = 0; k < len; ++k) codes[k] = (long long)(packed[k] >> 16) * n + (packed[k] & 0xffff);
  std::sort(codes.begin(), codes.end());
  for (int k = 0; k < len; ++k) out[k] = codes[k];
  return AWSEM_OK;
}

int awsem_batch_last_error(const awsem_batch *b, char *buf, size_t n)
{
  if (!b || !buf || n == 0) return AWSEM_E_ARG;
  std::snprintf(buf, n, "%s", b->err.c_str());
  return AWSEM_OK;
}

int awsem_batch_destroy(awsem_batch *b)
{
  delete b;
  return AWSEM_OK;
}

}  // extern "C"
