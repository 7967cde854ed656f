// awsem_devsys.cpp -- one system's device image: layout, packing and upload.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "awsem_host.h"

#define DS_HIPCHK(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return AwStatus::err(-7, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

AwStatus upload_raw(void **dst, const void *src, size_t bytes)
{
  DS_HIPCHK(hipMalloc(dst, std::max<size_t>(bytes, 16)));
  if (bytes) DS_HIPCHK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
  return AwStatus();
}

AwDevSystem::~AwDevSystem()
{
  if (arena) (void)hipFree(arena);
}

static size_t take(size_t &cur, size_t bytes)
{
  size_t o = cur;
  cur += (bytes + 255) & ~(size_t)255;
  return o;
}

template <class T>
static void put(std::vector<unsigned char> &img, size_t off, const std::vector<T> &v)
{
  if (!v.empty()) std::memcpy(img.data() + off, v.data(), v.size() * sizeof(T));
}

AwStatus aw_devsys_plan(const AwSystem &S, const FFParams *P, bool fp32, bool scratch, bool own_state, AwDevPlan &plan,
                        bool lib)
{
  plan = AwDevPlan();
  plan.fp32 = fp32; plan.scratch = scratch; plan.own_state = own_state;
  const FFParams &FP = P ? *P : S.P;
  if (FP.fm_flag && lib) {
    AwStatus s = aw_build_fm(S, FP, plan.fm);
    if (!s.ok()) return s;
  }
  int n = S.n;
  long long allp = (long long)n * (n - 1) / 2;
  plan.nl_cap = (int)std::min<long long>(std::max<long long>(allp, 1), 256LL * n);
  size_t tsz = fp32 ? sizeof(float) : sizeof(double);
  size_t cur = 0;
  if (own_state) {
    plan.off_x = take(cur, 9 * (size_t)n * sizeof(double));
    plan.off_v = take(cur, 9 * (size_t)n * sizeof(double));
    plan.off_f = take(cur, 9 * (size_t)n * sizeof(double));
    plan.off_e = take(cur, AW_N_ET * sizeof(double));
  }
  plan.off_rinfo = take(cur, (size_t)n * sizeof(int));
  plan.off_aps = take(cur, 12 * (size_t)n * sizeof(double));
  plan.off_chs = take(cur, (size_t)S.nch * sizeof(int));
  plan.off_chl = take(cur, (size_t)S.nch * sizeof(int));
  plan.off_items = take(cur, (size_t)plan.fm.n_items() * (fp32 ? sizeof(FMItem) : sizeof(FMItemD)));
  plan.off_rf = take(cur, plan.fm.rf.size() * tsz);
  plan.off_coef = take(cur, plan.fm.coef.size() * tsz);
  // table mode: items and curves (inside the fragment-memory range that
  // identical replicas share), curves as (energy, force) pairs in T
  std::vector<int> tb_items;
  if (FP.fm_tb_flag && lib) aw_fm_table_items(S, tb_items);
  plan.n_tb_items = (int)(tb_items.size() / 4);
  plan.off_tbi = take(cur, tb_items.size() * sizeof(int));
  plan.off_tbc = take(cur, lib ? 2 * S.fmt.e.size() * tsz : 0);
  plan.off_nl = take(cur, (AW_NL_HDR + 6 * (size_t)plan.nl_cap) * sizeof(int));
  plan.off_nlx = take(cur, 4 * (size_t)n * sizeof(double));
  // cell-list table for large systems (the wide engine builds their lists
  // through it): buckets = a power of two >= 2n
  int cell_min = AW_CELL_MIN_N;
  if (const char *e = std::getenv("AWSEM_CELL_MIN_N")) cell_min = std::atoi(e);
  if (n > cell_min) {
    int h = 1;
    while (h < 2 * n) h <<= 1;
    plan.cell_h = h;
    plan.off_cells = take(cur, (2 * (size_t)h + 4 * (size_t)n) * sizeof(int));
  }
  if (scratch) plan.off_scr = take(cur, (size_t)aw_slab_doubles(n) * sizeof(double));
  plan.off_fault = take(cur, 256);
  // coeff-toggled terms (per image: they are small and follow the
  // system's own side files)
  AwXtPack xt;
  if (FP.xt_on) {
    AwStatus s = aw_build_extras(S, FP, xt);
    if (!s.ok()) return s;
  }
  plan.has_go = !xt.go_start.empty();
  plan.has_zres = FP.memb_flag && !S.xt.zres.empty();
  plan.n_go = (int)xt.go.size(); plan.n_cr = (int)xt.cr.size(); plan.n_dh = (int)xt.dh.size(); plan.n_vfm = (int)xt.vfm.size();
  plan.off_gos = take(cur, xt.go_start.size() * sizeof(int));
  plan.off_go = take(cur, xt.go.size() * sizeof(XtGo));
  plan.off_cr = take(cur, xt.cr.size() * sizeof(XtPair));
  plan.off_dh = take(cur, xt.dh.size() * sizeof(XtPair));
  plan.off_vfm = take(cur, xt.vfm.size() * sizeof(XtPair));
  plan.off_zres = take(cur, plan.has_zres ? (size_t)n * sizeof(int) : 0);
  plan.go_norm = S.xt.go_norm;
  // the system's own reach (its contact list may widen pair_list_cutoff)
  // with the force field's list skin
  double lr = S.P.pair_list_cutoff + FP.list_skin;
  plan.list_r2 = lr * lr;
  plan.cr_cut2 = S.P.cr_glob_cutoff_sq;
  plan.bytes = std::max<size_t>(cur, 256);
  // the image: zeros (state, list headers, slab) and the packed constants
  std::vector<unsigned char> &img = plan.image;
  img.assign(plan.bytes, 0);
  put(img, plan.off_rinfo, S.rinfo);
  put(img, plan.off_aps, S.aps);
  put(img, plan.off_chs, S.ch_start);
  put(img, plan.off_chl, S.ch_len);
  const AwFMPack &fm = plan.fm;
  if (fp32) {
    std::vector<FMItem> items;
    std::vector<float> rf, coef;
    aw_fm_device(fm, items, rf, coef);
    put(img, plan.off_items, items);
    put(img, plan.off_rf, rf);
    put(img, plan.off_coef, coef);
  } else {
    std::vector<FMItemD> items;
    std::vector<double> rf, coef;
    aw_fm_device(fm, items, rf, coef);
    put(img, plan.off_items, items);
    put(img, plan.off_rf, rf);
    put(img, plan.off_coef, coef);
  }
  put(img, plan.off_tbi, tb_items);
  put(img, plan.off_gos, xt.go_start);
  put(img, plan.off_go, xt.go);
  put(img, plan.off_cr, xt.cr);
  put(img, plan.off_dh, xt.dh);
  put(img, plan.off_vfm, xt.vfm);
  if (plan.has_zres) put(img, plan.off_zres, S.xt.zres);
  if (lib && !S.fmt.e.empty()) {
    const AwFMTable &T = S.fmt;
    if (fp32) {
      std::vector<float> c(2 * T.e.size());
      for (size_t q = 0; q < T.e.size(); ++q) { c[2 * q] = (float)T.e[q]; c[2 * q + 1] = (float)T.f[q]; }
      put(img, plan.off_tbc, c);
    } else {
      std::vector<double> c(2 * T.e.size());
      for (size_t q = 0; q < T.e.size(); ++q) { c[2 * q] = T.e[q]; c[2 * q + 1] = T.f[q]; }
      put(img, plan.off_tbc, c);
    }
  }
  plan.n_items = fm.n_items();
  plan.fm_ij.resize(plan.n_items);
  for (int k = 0; k < plan.n_items; ++k) plan.fm_ij[k] = fm.meta[5 * k];
  plan.n = n;
  plan.nch = S.nch;
  plan.fm = AwFMPack();  // packed into the image
  return AwStatus();
}


AwStatus aw_devsys_create(const AwDevPlan &plan, AwDevSystem &out)
{
  int n = plan.n;
  if (plan.image.size() != plan.bytes) return AwStatus::err(-10, "device image already consumed");
  DS_HIPCHK(hipMalloc(&out.arena, plan.bytes));
  out.bytes = plan.bytes;
  DS_HIPCHK(hipMemcpy(out.arena, plan.image.data(), plan.bytes, hipMemcpyHostToDevice));
  unsigned char *b = (unsigned char *)out.arena;
  ReplicaDesc &R = out.d;
  std::memset(&R, 0, sizeof R);
  R.n = n; R.nch = plan.nch;
  R.fm_items = plan.n_items;
  R.nl_cap = plan.nl_cap;
  if (plan.own_state) {
    R.x = (double *)(b + plan.off_x);
    R.v = (double *)(b + plan.off_v);
    R.f = (double *)(b + plan.off_f);
    R.energy = (double *)(b + plan.off_e);
  }
  R.rinfo = (const int *)(b + plan.off_rinfo);
  R.aps = (const double *)(b + plan.off_aps);
  R.ch_start = (const int *)(b + plan.off_chs);
  R.ch_len = (const int *)(b + plan.off_chl);
  R.items = b + plan.off_items;
  R.rf = b + plan.off_rf;
  R.coef = b + plan.off_coef;
  R.nl = (int *)(b + plan.off_nl);
  R.nlx = (double *)(b + plan.off_nlx);
  R.scratch = plan.scratch ? (double *)(b + plan.off_scr) : nullptr;
  R.cells = plan.cell_h ? (int *)(b + plan.off_cells) : nullptr;
  R.tb_items = plan.n_tb_items;
  R.tb_item = (const int *)(b + plan.off_tbi);
  R.tb_curve = b + plan.off_tbc;
  R.fault = (int *)(b + plan.off_fault);
  R.cell_h = plan.cell_h;
  R.n_go = plan.n_go; R.n_cr = plan.n_cr; R.n_dh = plan.n_dh; R.n_vfm = plan.n_vfm;
  R.go_start = plan.has_go ? (const int *)(b + plan.off_gos) : nullptr;
  R.go_item = b + plan.off_go;
  R.cr_item = b + plan.off_cr;
  R.dh_item = b + plan.off_dh;
  R.vfm_item = b + plan.off_vfm;
  R.zres = plan.has_zres ? (const int *)(b + plan.off_zres) : nullptr;
  R.go_norm = plan.go_norm;
  R.list_r2 = plan.list_r2;
  R.cr_cut2 = plan.cr_cut2;
  return AwStatus();
}
