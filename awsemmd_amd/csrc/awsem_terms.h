// awsem_terms.h -- the FixBackbone force terms and the Langevin update as
// per-item functions over one replica view.  Included by the gfx950 kernel
// file; every function is also valid host C++ (AW_HOST_EMU), which lets
// the exact kernel arithmetic be stepped through on a CPU when debugging.
#pragma once
#include <type_traits>
#include <stdint.h>
#include <math.h>
#include "awsem_types.h"

#ifdef AW_HOST_EMU
#include <algorithm>
#include <string.h>
#define AW_DEV inline
#define AW_DEVCONST static const
struct int2 { int x, y; };
struct float4 { float x, y, z, w; };
struct double4 { double x, y, z, w; };
inline float aw_fast_exp(float a) { return expf(a); }
inline float aw_fast_exp2(float a) { return exp2f(a); }
inline float aw_fsqrt(float a) { return sqrtf(a); }
inline float aw_frcp(float a) { return 1.0f / a; }
inline float aw_frsq(float a) { return 1.0f / sqrtf(a); }
inline void aadd(double *p, double v) { *p += v; }
inline void amax_i(int *p, int v) { if (v > *p) *p = v; }
inline int aadd_i(int *p, int v) { int o = *p; *p += v; return o; }
inline int f2i(float f) { int i; memcpy(&i, &f, 4); return i; }
inline float i2f(int i) { float f; memcpy(&f, &i, 4); return f; }
#define AW_BARRIER() ((void)0)
#define AW_NOINLINE inline
#define AW_HOSTDEV
typedef const FFParams *ParamPtr;
#define AW_SCHED_FENCE() ((void)0)
#define AW_G
inline void aw_hold(int &) {}
inline double aw_sval(double v) { return v; }
template <bool L> inline void aadd_as(double *p, double v) { *p += v; }
template <bool L> struct aw_as { typedef double type; };
template <bool L> inline double *aw_cast(double *p) { return p; }
template <bool L> inline const double *aw_cast(const double *p) { return p; }
inline void aadd_t(double *p, double v) { *p += v; }
template <class ItemT> inline ItemT ld_item(const ItemT *p) { return *p; }
template <class P> inline P *aw_uniform(P *p) { return p; }
inline unsigned aw_umul24(unsigned a, unsigned b) { return (a & 0xffffffu) * (b & 0xffffffu); }
// max of non-negative floats into *p (as int bits)
inline void wave_amax(int *p, float v) { amax_i(p, f2i(v)); }
// wave-level queue reservation (one lane here): a lane's ballot is its own
// predicate and it ranks first among the lanes that set it
typedef unsigned long long lmask;
inline lmask w_ballot(bool p) { return p ? 1ull : 0ull; }
inline int w_count(lmask m) { return (int)m; }
inline int w_rank(lmask) { return 0; }
inline void q_reserve4(int *ctr, const int n[4], int b[4])
{
  for (int k = 0; k < 4; ++k) { b[k] = ctr[k]; ctr[k] += n[k]; }
}
inline int q_reserve1(int *ctr, int n) { int b = *ctr; *ctr += n; return b; }
#else
#include <hip/hip_runtime.h>
#define AW_DEV __device__ __forceinline__
#define AW_DEVCONST __device__ __constant__
__device__ __forceinline__ float aw_fast_exp(float a) { return __expf(a); }
__device__ __forceinline__ float aw_fast_exp2(float a) { return __builtin_amdgcn_exp2f(a); }
// the hardware square root: the same value as the rounded form for every
// normal input (that form only rescales inputs below 2^-96 around it); its
// one use is an interatomic distance
__device__ __forceinline__ float aw_fsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ float aw_frcp(float a) { return __builtin_amdgcn_rcpf(a); }
__device__ __forceinline__ float aw_frsq(float a) { return __builtin_amdgcn_rsqf(a); }
__device__ __forceinline__ void aadd(double *p, double v) { atomicAdd(p, v); }
__device__ __forceinline__ void amax_i(int *p, int v) { atomicMax(p, v); }
__device__ __forceinline__ int aadd_i(int *p, int v) { return atomicAdd(p, v); }
__device__ __forceinline__ int f2i(float f) { return __float_as_int(f); }
__device__ __forceinline__ float i2f(int i) { return __int_as_float(i); }
#define AW_BARRIER() __syncthreads()
// a separately allocated device function (called once per pass)
#define AW_NOINLINE __device__ __attribute__((noinline))
#define AW_HOSTDEV __host__ __device__
// the force-field parameters through the constant address space: loads
// from it are known invariant, so the compiler hoists them out of the pair
// loops (a plain pointer in a struct carries no aliasing guarantee, and
// every atomic in a loop would force them to be reloaded)
typedef const __attribute__((address_space(4))) FFParams *ParamPtr;
// instruction-scheduling barrier: nothing is moved across it
#define AW_SCHED_FENCE() __builtin_amdgcn_sched_barrier(0)
// global-memory pointee: global_ instructions instead of flat ones (a flat
// access also counts against the LDS wait counter, so every LDS wait after
// it would wait for the global round trip too)
#define AW_G __attribute__((address_space(1)))
// a no-op that reads x: the wait for a load into x is placed here, not at
// x's first real use (only a register operand; nothing is loaded)
__device__ __forceinline__ void aw_hold(int &x) { asm volatile("" : "+v"(x)); }
// a wave-uniform value kept in scalar registers: a per-lane choice between
// two such values then stays a choice, instead of being folded into one
// load from a per-lane address (a vector load and its wait)
__device__ __forceinline__ double aw_sval(double v)
{
  unsigned long long b = __double_as_longlong(v);
  unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b), hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// typed accesses for code compiled apart from the kernel (a noinline
// function), where a plain pointer would give flat instructions: aadd_as
// adds into LDS (L) or global memory, ld_item loads a 16-byte multiple
// struct from global memory in 16-byte vectors
typedef __attribute__((address_space(3))) double aw_lds_double;
typedef __attribute__((address_space(1))) double aw_global_double;
template <bool L>
__device__ __forceinline__ void aadd_as(double *p, double v)
{
  if (L)
    __hip_atomic_fetch_add((aw_lds_double *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else
    __hip_atomic_fetch_add((aw_global_double *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// a base pointer converted to its address space once (aw_cast): indexing it
// then needs no per-access conversion (a generic-to-LDS conversion carries a
// null check, and generic addresses are 64-bit)
template <bool L> struct aw_as { typedef aw_global_double type; };
template <> struct aw_as<true> { typedef aw_lds_double type; };
template <bool L>
__device__ __forceinline__ typename aw_as<L>::type *aw_cast(double *p) { return (typename aw_as<L>::type *)p; }
template <bool L>
__device__ __forceinline__ const typename aw_as<L>::type *aw_cast(const double *p)
{
  return (const typename aw_as<L>::type *)p;
}
__device__ __forceinline__ void aadd_t(aw_lds_double *p, double v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void aadd_t(aw_global_double *p, double v)
{
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// a wave-uniform pointer moved to scalar registers (arguments of a device
// function arrive in vector registers)
template <class P>
__device__ __forceinline__ P *aw_uniform(P *p)
{
  unsigned long long v = (unsigned long long)p;
  unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v), hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return (P *)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ unsigned aw_umul24(unsigned a, unsigned b) { return __umul24(a, b); }
template <class ItemT>
__device__ __forceinline__ ItemT ld_item(const ItemT *p)
{
  static_assert(sizeof(ItemT) % 16 == 0, "item size");
  typedef int I4 __attribute__((ext_vector_type(4)));
  const __attribute__((address_space(1))) I4 *q = (const __attribute__((address_space(1))) I4 *)p;
  I4 w[sizeof(ItemT) / 16];
#pragma unroll
  for (int u = 0; u < (int)(sizeof(ItemT) / 16); ++u) w[u] = q[u];
  ItemT it;
  __builtin_memcpy(&it, w, sizeof it);
  return it;
}
// max of non-negative floats into *p (as int bits): reduced across the
// wave first (every lane active), then one atomic per wave
__device__ __forceinline__ void wave_amax(int *p, float v)
{
  int b = __float_as_int(v);
  for (int o = 32; o > 0; o >>= 1) {
    int u = __shfl_xor(b, o);
    b = u > b ? u : b;
  }
  if (__lane_id() == 0) atomicMax(p, b);
}
// wave-level work-queue reservation: w_ballot / w_count over the active
// lanes, w_rank = the number of lanes below this one in a ballot
typedef unsigned long long lmask;
__device__ __forceinline__ lmask w_ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int w_count(lmask m) { return __popcll(m); }
__device__ __forceinline__ int w_rank(lmask m) { return __popcll(m & ((1ull << __lane_id()) - 1ull)); }
// reserves n[k] (wave-uniform) slots in each of four queues whose counters
// ctr[0..3] form two 8-byte-aligned 64-bit words (counts stay below 2^31,
// so the low half never carries): both atomics are issued by the first
// active lane back to back, and the bases come back through readfirstlane
// (no LDS round trip for the broadcast)
__device__ __forceinline__ void q_reserve4(int *ctr, const int n[4], int b[4])
{
  int first = __ffsll((unsigned long long)__ballot(1)) - 1;
  unsigned long long o0 = 0, o1 = 0;
  if (__lane_id() == first) {
    unsigned long long *c = (unsigned long long *)ctr;
    o0 = atomicAdd(c, (unsigned long long)(unsigned)n[0] | ((unsigned long long)(unsigned)n[1] << 32));
    o1 = atomicAdd(c + 1, (unsigned long long)(unsigned)n[2] | ((unsigned long long)(unsigned)n[3] << 32));
  }
  b[0] = __builtin_amdgcn_readfirstlane((int)(unsigned)o0);
  b[1] = __builtin_amdgcn_readfirstlane((int)(unsigned)(o0 >> 32));
  b[2] = __builtin_amdgcn_readfirstlane((int)(unsigned)o1);
  b[3] = __builtin_amdgcn_readfirstlane((int)(unsigned)(o1 >> 32));
}
// n (wave-uniform) slots of one queue: one atomic per wave
__device__ __forceinline__ int q_reserve1(int *ctr, int n)
{
  int first = __ffsll((unsigned long long)__ballot(1)) - 1;
  int o = 0;
  if (__lane_id() == first) o = atomicAdd(ctr, n);
  return __builtin_amdgcn_readfirstlane(o);
}
#endif

#define AW_PAP_DELTA 1e-12
#define AW_DSSP_NU_DELTA 1e-4
#define AW_DELTA 0.00001
#define AW_DELTA_WATER_XI 1e-8
#define AW_DELTA_HELIX_XI 1e-6

// ------------------------------------------------------------------------
// small math helpers

template <class T> struct V3 { T x, y, z; };

AW_DEV float t_exp(float a) { return aw_fast_exp(a); }
AW_DEV double t_exp(double a) { return exp(a); }
AW_DEV float t_exp2(float a) { return aw_fast_exp2(a); }
AW_DEV double t_exp2(double a) { return exp2(a); }
AW_DEV float t_sqrt(float a) { return aw_fsqrt(a); }
// sqrt of an FP64 value that no predicate depends on: FP64 exact, or (FP32
// path) the FP32 reciprocal square root refined by one FP64 Newton step
// (relative error ~1e-13: enough for the FP64-formed offsets and switching
// arguments it feeds)
template <class T>
AW_DEV double sqrt_t(double x)
{
  if (sizeof(T) == 8) return sqrt(x);
  double y = (double)aw_frsq(fmaxf((float)x, 1e-30f));
  y = y * (1.5 - 0.5 * x * y * y);
  return x * y;
}
// a / b: FP32 through the hardware reciprocal (1 ulp), FP64 exact
AW_DEV float t_div(float a, float b) { return a * aw_frcp(b); }
AW_DEV double t_div(double a, double b) { return a / b; }
AW_DEV double t_sqrt(double a) { return sqrt(a); }
AW_DEV float t_tanh(float a) { return tanhf(a); }
AW_DEV double t_tanh(double a) { return tanh(a); }
AW_DEV float t_cos(float a) { return cosf(a); }
AW_DEV double t_cos(double a) { return cos(a); }
AW_DEV float t_sin(float a) { return sinf(a); }
AW_DEV double t_sin(double a) { return sin(a); }
AW_DEV float t_atan2(float y, float x) { return atan2f(y, x); }
AW_DEV double t_atan2(double y, double x) { return atan2(y, x); }

// exact (no FMA contraction) squared distance: the list predicates must
// decide exactly as the reference's FP64 host code does
AW_DEV double rsq_exact(const double *a, const double *b)
{
#pragma clang fp contract(off)
  double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
  return d0 * d0 + d1 * d1 + d2 * d2;
}

// FP32 estimate of |a - b|^2 (relative error below 1e-6): predicates on
// rsq_exact are decided with it unless it lies within AW_RSQ_MARGIN of the
// threshold, where the exact FP64 value decides
#define AW_RSQ_MARGIN 1e-5f
AW_DEV float rsq_f(const double *a, const double *b)
{
  float d0 = (float)(a[0] - b[0]), d1 = (float)(a[1] - b[1]), d2 = (float)(a[2] - b[2]);
  return d0 * d0 + d1 * d1 + d2 * d2;
}

// exact (rsq_exact(a, b) < t2) for t2 > 0, given e = rsq_f(a, b)
AW_DEV bool rsq_lt_est(float e, const double *a, const double *b, double t2)
{
  float tf = (float)t2;
  if (e < tf * (1.0f - AW_RSQ_MARGIN)) return true;
  if (e > tf * (1.0f + AW_RSQ_MARGIN)) return false;
  return rsq_exact(a, b) < t2;
}
AW_DEV bool rsq_lt(const double *a, const double *b, double t2) { return rsq_lt_est(rsq_f(a, b), a, b, t2); }
// exact (rsq_exact(a, b) < t2 && !(sqrt(rsq_exact(a, b)) > t)) for t2 = t * t
// rounded: well inside (outside) the FP32 estimate decides, since then
// r^2 < t2 (1 - 9e-6) implies sqrt(r^2) < t
AW_DEV bool dist_within(const double *a, const double *b, double t, double t2)
{
  float e = rsq_f(a, b), tf = (float)t2;
  if (e < tf * (1.0f - AW_RSQ_MARGIN)) return true;
  if (e > tf * (1.0f + AW_RSQ_MARGIN)) return false;
  double r2 = rsq_exact(a, b);
  return r2 < t2 && !(sqrt(r2) > t);
}

AW_DEV double lin3_exact(double a, double xa, double b, double xb, double c, double xc)
{
#pragma clang fp contract(off)
  return a * xa + b * xb + c * xc;
}

template <class T>
AW_DEV V3<T> dvec(const double *a, const double *b)
{
  V3<T> d;
  d.x = (T)(a[0] - b[0]);
  d.y = (T)(a[1] - b[1]);
  d.z = (T)(a[2] - b[2]);
  return d;
}
template <class T> AW_DEV T dot(const V3<T> &a, const V3<T> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> AW_DEV V3<T> crs(const V3<T> &a, const V3<T> &b)
{
  V3<T> c;
  c.x = a.y * b.z - a.z * b.y;
  c.y = a.z * b.x - a.x * b.z;
  c.z = a.x * b.y - a.y * b.x;
  return c;
}


// ------------------------------------------------------------------------
// per-replica view: every pointer is either LDS or a global scratch slab

#define AW_SC_STRIDE 11  // doubles per residue for the ten per-residue scalars

template <class T>
struct View {
  int n, nch;
  const int *rinfo;     // [n] (global, or its copy in LDS)
  const AW_G double *aps;    // [n][12]
  const AW_G int *ch_start;  // [nch] 0-based
  const AW_G int *ch_len;
  double *s;            // working slab (LDS or global): see the accessors
  // where the pair terms add forces (FA): atom (kind, i) at
  // fb + 3 (kind fs + i).  The slab's force array (fb = f(), fs = n), or a
  // workgroup's LDS window holding only the kinds a pass writes
  double *fb;
  int fs;
  double *ered;         // energy accumulators [N_ERED] (LDS)
  ParamPtr P;
  // fragment memory
  int fm_items, fm_item_base;
  // fragment-memory table mode: items and curves (ReplicaDesc), error word
  int tb_items;
  const AW_G int *tb_item;
  const void *tb_curve;
  AW_G int *fault;
  const T *coef;
  const void *rf;        // T4 per (memory, pair)
  const void *items;     // FMItem / FMItemD per item
  // Verlet list of residue pairs (resident engine; nl == nullptr: sweep all pairs)
  AW_G int *nl;          // global: header (AW_NL_HDR ints: [0] count, [1] valid, [2] builds), then (i<<16 | j)
  int nl_cap;
  AW_G double *nlx;      // global: CA positions [n][3] and extents [n] at the last build
  int *nl_state;         // LDS: [0] max drift (float bits), [1] build counter, [2..5] queue counts,
                         // [6] largest extent, [7] list pairs + 1 when valid, else 0
  AW_G int *wq;          // global work queues beside the list: DSSP [2 cap], P_AP [cap], EV [cap], density pairs [cap]
  // fragment-memory sharing (resident engine; sh_n == 0 elsewhere): the
  // replica's claim word and chunk table, and (LDS) the chunks this
  // workgroup claimed in the current evaluation
  int sh_n;
  unsigned long long *sh_word;
  const ShareChunk *sh_chunks;
  unsigned long long *sh_mask;
  // slab layout: x[3][n][3] (CA, CB, O), xv[3][n][3] (N, H, C'),
  // f[3][n][3], the ten per-residue scalars [n][AW_SC_STRIDE], ext (float)
  AW_DEV double *x() const { return s; }
  AW_DEV double *xv() const { return s + 9 * n; }
  AW_DEV double *f() const { return s + 18 * n; }
  // the per-residue scalars are stored residue by residue: one base
  // address serves all ten (the scalar's index is an instruction offset),
  // which frees the registers ten array bases took in the pair loops; the
  // stride of 11 doubles keeps 16 consecutive residues on distinct LDS banks
  struct SRow {
    double *p;
    AW_DEV double &operator[](int i) const { return p[AW_SC_STRIDE * i]; }
  };
  AW_DEV SRow sc(int k) const { return SRow{s + 27 * n + k}; }
  AW_DEV SRow rho_w() const { return sc(0); }
  AW_DEV SRow rho_h() const { return sc(1); }
  AW_DEV SRow sig_w() const { return sc(2); }
  AW_DEV SRow sigp_w() const { return sc(3); }
  AW_DEV SRow sig_h() const { return sc(4); }
  AW_DEV SRow sigp_h() const { return sc(5); }
  AW_DEV SRow xi_w() const { return sc(6); }
  AW_DEV SRow bf() const { return sc(7); }
  AW_DEV SRow hxi1() const { return sc(8); }
  AW_DEV SRow hxi2() const { return sc(9); }
  AW_DEV float *ext() const { return (float *)(s + (27 + AW_SC_STRIDE) * n); }
  // resident engine: a copy of rinfo after the extents
  AW_DEV int *rinfo_copy() const { return (int *)(s + (27 + AW_SC_STRIDE) * n + (n + 1) / 2); }
};

#define XA(V, kind, i) ((V).x() + ((kind) * (V).n + (i)) * 3)
#define XV(V, kind, i) ((V).xv() + ((kind) * (V).n + (i)) * 3)
#define FA(V, kind, i) ((V).fb + ((kind) * (V).fs + (i)) * 3)
enum { K_CA = 0, K_CB = 1, K_O = 2 };
enum { K_N = 0, K_H = 1, K_CP = 2 };

template <class T, class R>
AW_DEV void fadd(const View<T> &V, int kind, int i, R s, const V3<R> &d)
{
  double *p = FA(V, kind, i);
  aadd(p + 0, (double)(s * d.x));
  aadd(p + 1, (double)(s * d.y));
  aadd(p + 2, (double)(s * d.z));
}
template <class T, class R>
AW_DEV void fadd3(const View<T> &V, int kind, int i, R fx, R fy, R fz)
{
  double *p = FA(V, kind, i);
  aadd(p + 0, (double)fx);
  aadd(p + 1, (double)fy);
  aadd(p + 2, (double)fz);
}

// Forces of one residue's local terms (chain, chi, shake, Ramachandran,
// bonds, the CB-O excluded volume): every one of them acts on CA(i-1),
// O(i-1), CA(i), O(i), CB(i) and CA(i+1) only, so they are summed in
// registers and added to the force array once per residue (18 atomic adds
// instead of ~130).  add(kind, d, ...) names the atom by its residue
// offset d in {-1, 0, 1}; both are constants wherever the terms are inlined.
struct LocF {
  double f[6][3];
  AW_DEV void clear()
  {
#pragma unroll
    for (int a = 0; a < 6; ++a) f[a][0] = f[a][1] = f[a][2] = 0.0;
  }
  AW_DEV static int slot(int kind, int d)
  {
    return d < 0 ? (kind == K_CA ? 0 : 1) : (d > 0 ? 5 : (kind == K_CA ? 2 : (kind == K_O ? 3 : 4)));
  }
  template <class R>
  AW_DEV void add3(int kind, int d, R fx, R fy, R fz)
  {
    double *q = f[slot(kind, d)];
    q[0] += (double)fx; q[1] += (double)fy; q[2] += (double)fz;
  }
  template <class R>
  AW_DEV void add(int kind, int d, R s, const V3<R> &v) { add3(kind, d, s * v.x, s * v.y, s * v.z); }
  // CA(i-1) and O(i-1) exist unless i is a chain's first residue, CA(i+1)
  // unless it is the last
  template <class T>
  AW_DEV void flush(const View<T> &V, int i, bool first, bool last) const
  {
    if (!first) { fadd3(V, K_CA, i - 1, f[0][0], f[0][1], f[0][2]); fadd3(V, K_O, i - 1, f[1][0], f[1][1], f[1][2]); }
    fadd3(V, K_CA, i, f[2][0], f[2][1], f[2][2]);
    fadd3(V, K_O, i, f[3][0], f[3][1], f[3][2]);
    fadd3(V, K_CB, i, f[4][0], f[4][1], f[4][2]);
    if (!last) fadd3(V, K_CA, i + 1, f[5][0], f[5][1], f[5][2]);
  }
};

// energy accumulators (per thread, FP64)
struct EAcc {
  double chain, chi, rama, ev, dssp, pap, water, burial, helix, fm, bond, shake;
  AW_DEV void clear() { chain = chi = rama = ev = dssp = pap = water = burial = helix = fm = bond = shake = 0.0; }
};
#define N_EACC 12
// energy accumulators of a workgroup: EAcc's twelve, then the six
// coeff-toggled terms (awsem_extra.h)
#define N_ERED 18
AW_DEV double eacc_get(const EAcc &e, int k)
{
  switch (k) {
  case 0: return e.chain; case 1: return e.chi; case 2: return e.rama; case 3: return e.ev;
  case 4: return e.dssp; case 5: return e.pap; case 6: return e.water; case 7: return e.burial;
  case 8: return e.helix; case 9: return e.fm; case 10: return e.bond; default: return e.shake;
  }
}
AW_DEVCONST int eacc_slot[N_ERED] = {AW_ET_CHAIN, AW_ET_CHI, AW_ET_RAMA, AW_ET_VEXCLUDED, AW_ET_DSSP,
                                                 AW_ET_PAP, AW_ET_WATER, AW_ET_BURIAL, AW_ET_HELIX, AW_ET_FRAGMEM,
                                                 AW_ET_BOND, AW_ET_SHAKE, AW_ET_AMHGO, AW_ET_VFRAGMEM,
                                                 AW_ET_CONT_REST, AW_ET_MEMB, AW_ET_SSB, AW_ET_DH};

// adds a thread's energy accumulators into ered[N_EACC] (LDS): a wave
// reduction, then one LDS atomic per wave and term
#ifdef AW_HOST_EMU
AW_DEV void eflush(const EAcc &E, double *ered)
{
  for (int k = 0; k < N_EACC; ++k) ered[k] += eacc_get(E, k);
}
#else
AW_DEV double wave_sum_d(double v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
AW_DEV void eflush(const EAcc &E, double *ered)
{
  int lane = threadIdx.x & 63;
  for (int k = 0; k < N_EACC; ++k) {
    double v = wave_sum_d(eacc_get(E, k));
    if (lane == 0 && v != 0.0) atomicAdd(&ered[k], v);
  }
}
#endif

#define ON(P, slot) ((P)->term_mask & (1u << (slot)))

AW_DEV int rtype(int ri) { return ri & RI_TYPE_MASK; }
AW_DEV int rchain(int ri) { return ri >> RI_CHAIN_SHIFT; }

// unordered residue pair k -> (i, j), i < j (each pair exactly once)
AW_DEV void pair_decode(int k, int n, int &i, int &j)
{
  int H = (n - 1) >> 1;
  int a, b;
  if (k < n * H) {
    a = k % n;
    int d = 1 + k / n;
    b = a + d;
    if (b >= n) b -= n;
  } else {
    a = k - n * H;
    b = a + (n >> 1);
  }
  i = a < b ? a : b;
  j = a < b ? b : a;
}
AW_DEV int pair_count(int n) { return n * (n - 1) / 2; }

// ------------------------------------------------------------------------
// Phase A: virtual atoms (fb.cpp:7113-7138), extents, zeroing

template <class T>
AW_DEV void phase_A(const View<T> &V, int t0, int ts)
{
  ParamPtr P = V.P;
  int n = V.n;
  float dmax = 0.0f, emax = 0.0f;  // this lane's largest drift and extent
  for (int i = t0; i < n; i += ts) {
    int ri = V.rinfo[i];
    bool first = ri & RI_FIRST, last = ri & RI_LAST;
    // every position the residue needs, read into registers before any
    // store (the stores into the slab could alias the reads otherwise, and
    // each read would wait for the previous store); neighbour indices are
    // clamped, their values unused at a chain's ends
    int im = first ? i : i - 1, ip = last ? i : i + 1;
    double ca[3], o[3], cb[3], cam[3], om[3], cap[3], c0[3] = {0, 0, 0}, e0 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      ca[k] = XA(V, K_CA, i)[k]; o[k] = XA(V, K_O, i)[k]; cb[k] = XA(V, K_CB, i)[k];
      cam[k] = XA(V, K_CA, im)[k]; om[k] = XA(V, K_O, im)[k]; cap[k] = XA(V, K_CA, ip)[k];
    }
    if (V.nl) {
      c0[0] = V.nlx[3 * i]; c0[1] = V.nlx[3 * i + 1]; c0[2] = V.nlx[3 * i + 2];
      e0 = V.nlx[3 * n + i];
    }
    double xn[3], xh[3], xcp[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      xn[k] = first ? 0.0 : lin3_exact(P->an, cam[k], P->bn, ca[k], P->cn, om[k]);
      xh[k] = first ? 0.0 : lin3_exact(P->ah, cam[k], P->bh, ca[k], P->ch, om[k]);
      xcp[k] = last ? 0.0 : lin3_exact(P->ap, ca[k], P->bp, cap[k], P->cp, o[k]);
      XV(V, K_N, i)[k] = xn[k]; XV(V, K_H, i)[k] = xh[k]; XV(V, K_CP, i)[k] = xcp[k];
    }
    double e2 = rsq_exact(ca, cb);
    e2 = fmax(e2, rsq_exact(ca, o));
    if (!first) { e2 = fmax(e2, rsq_exact(ca, xn)); e2 = fmax(e2, rsq_exact(ca, xh)); }
    if (!last) e2 = fmax(e2, rsq_exact(ca, xcp));
    float ei = (float)(sqrt(e2) * (1.0 + 1e-6) + 1e-4);
    V.ext()[i] = ei;
    if (V.nl) {
      // CA drift plus growth of the extent since the last list build
      double d2 = rsq_exact(ca, c0);
      float drift = (float)sqrt(d2) * 1.0001f + 1e-4f + fmaxf(0.0f, ei - (float)e0);
      dmax = fmaxf(dmax, drift);
      emax = fmaxf(emax, ei);
    }
    double *f = V.f();
    for (int k = 0; k < 3; ++k) {
      f[(0 * n + i) * 3 + k] = 0.0;
      f[(1 * n + i) * 3 + k] = 0.0;
      f[(2 * n + i) * 3 + k] = 0.0;
    }
    V.rho_w()[i] = 0.0; V.rho_h()[i] = 0.0; V.xi_w()[i] = 0.0;
    V.hxi1()[i] = 0.0; V.hxi2()[i] = 0.0; V.bf()[i] = 0.0;
  }
  // every lane of the wave is back here: one atomic per wave for the
  // largest drift and the largest extent (cell size of a cell-list build)
  if (V.nl) {
    wave_amax(&V.nl_state[0], dmax);
    wave_amax(&V.nl_state[6], emax);
  }
}

// ------------------------------------------------------------------------
// local terms

// Local (per-residue) terms run in FP64 in both precisions: they are O(n)
// and the dihedral slopes lose accuracy in FP32 for near-collinear atoms.

// compute_chain_potential, fb.cpp:2062-2170
template <class T>
AW_DEV void chain_term(const View<T> &V, int i, int ri, EAcc &E, LocF &F)
{
  ParamPtr P = V.P;
  bool first = ri & RI_FIRST, last = ri & RI_LAST, gly = ri & RI_GLY;
  if (!first && !gly) {
    V3<double> d = dvec<double>(XV(V, K_N, i), XA(V, K_CB, i));
    double r = sqrt(dot(d, d)), dr = r - (double)P->r_ncb0;
    double force = (double)2 * (double)P->k_chain[0] * dr / r;
    E.chain += (double)((double)P->k_chain[0] * dr * dr);
    F.add(K_CA, -1, -(double)P->an * force, d);
    F.add(K_O, -1, -(double)P->cn * force, d);
    F.add(K_CA, 0, -(double)P->bn * force, d);
    F.add(K_CB, 0, force, d);
  }
  if (!last && !gly) {
    V3<double> d = dvec<double>(XV(V, K_CP, i), XA(V, K_CB, i));
    double r = sqrt(dot(d, d)), dr = r - (double)P->r_cpcb0;
    double force = (double)2 * (double)P->k_chain[1] * dr / r;
    E.chain += (double)((double)P->k_chain[1] * dr * dr);
    F.add(K_CA, 1, -(double)P->bp * force, d);
    F.add(K_CA, 0, -(double)P->ap * force, d);
    F.add(K_O, 0, -(double)P->cp * force, d);
    F.add(K_CB, 0, force, d);
  }
  if (!first && !last) {
    V3<double> d = dvec<double>(XV(V, K_N, i), XV(V, K_CP, i));
    double r = sqrt(dot(d, d)), dr = r - (double)P->r_ncp0;
    double force = (double)2 * (double)P->k_chain[2] * dr / r;
    E.chain += (double)((double)P->k_chain[2] * dr * dr);
    F.add(K_CA, -1, -(double)P->an * force, d);
    F.add(K_O, -1, -(double)P->cn * force, d);
    F.add(K_CA, 1, (double)P->bp * force, d);
    F.add(K_CA, 0, -(double)(P->bn - P->ap) * force, d);
    F.add(K_O, 0, (double)P->cp * force, d);
  }
}

// compute_shake, fb.cpp:2172-2237
template <class T>
AW_DEV void shake_term(const View<T> &V, int i, int ri, EAcc &E, LocF &F)
{
  ParamPtr P = V.P;
  double ks = (double)(P->epsilon * P->k_shake);
  bool last = ri & RI_LAST;
  if (!last) {
    V3<double> d = dvec<double>(XA(V, K_CA, i), XA(V, K_CA, i + 1));
    double r = sqrt(dot(d, d)), dr = r - (double)P->r_sh1, force = (double)2 * ks * dr / r;
    E.shake += (double)(ks * dr * dr);
    F.add(K_CA, 0, -force, d);
    F.add(K_CA, 1, force, d);
  }
  {
    V3<double> d = dvec<double>(XA(V, K_CA, i), XA(V, K_O, i));
    double r = sqrt(dot(d, d)), dr = r - (double)P->r_sh2, force = (double)2 * ks * dr / r;
    E.shake += (double)(ks * dr * dr);
    F.add(K_CA, 0, -force, d);
    F.add(K_O, 0, force, d);
  }
  if (!last) {
    V3<double> d = dvec<double>(XA(V, K_O, i), XA(V, K_CA, i + 1));
    double r = sqrt(dot(d, d)), dr = r - (double)P->r_sh3, force = (double)2 * ks * dr / r;
    E.shake += (double)(ks * dr * dr);
    F.add(K_O, 0, -force, d);
    F.add(K_CA, 1, force, d);
  }
}

// compute_chi_potential, fb.cpp:2239-2339
template <class T>
AW_DEV void chi_term(const View<T> &V, int i, EAcc &E, LocF &F)
{
  ParamPtr P = V.P;
  V3<double> a = dvec<double>(XV(V, K_CP, i), XA(V, K_CA, i));
  V3<double> b = dvec<double>(XA(V, K_CA, i), XV(V, K_N, i));
  V3<double> c = dvec<double>(XA(V, K_CA, i), XA(V, K_CB, i));
  double arvsq = (double)1 / dot(a, a), brvsq = (double)1 / dot(b, b), crvsq = (double)1 / dot(c, c);
  double norm = sqrt(arvsq * brvsq * crvsq);
  V3<double> axb = crs(a, b), cxa = crs(c, a), bxc = crs(b, c);
  double chi = dot(axb, c) * norm;
  V3<double> ap_, bp_, cp_;
  ap_.x = norm * bxc.x - arvsq * chi * a.x; ap_.y = norm * bxc.y - arvsq * chi * a.y; ap_.z = norm * bxc.z - arvsq * chi * a.z;
  bp_.x = norm * cxa.x - brvsq * chi * b.x; bp_.y = norm * cxa.y - brvsq * chi * b.y; bp_.z = norm * cxa.z - brvsq * chi * b.z;
  cp_.x = norm * axb.x - crvsq * chi * c.x; cp_.y = norm * axb.y - crvsq * chi * c.y; cp_.z = norm * axb.z - crvsq * chi * c.z;
  double dchi = chi - (double)P->chi0;
  double force = (double)2 * (double)P->k_chi * dchi;
  E.chi += (double)((double)P->k_chi * dchi * dchi);
  F.add(K_CA, -1, (double)P->an * force, bp_);
  F.add(K_O, -1, (double)P->cn * force, bp_);
  F.add(K_CA, 1, -(double)P->bp * force, ap_);
  double c1 = (double)(1 - P->bn), c2 = (double)(P->ap - 1);
  F.add3(K_CA, 0, -(cp_.x + c1 * bp_.x + c2 * ap_.x) * force, -(cp_.y + c1 * bp_.y + c2 * ap_.y) * force,
        -(cp_.z + c1 * bp_.z + c2 * ap_.z) * force);
  F.add(K_O, 0, -(double)P->cp * force, ap_);
  F.add(K_CB, 0, force, cp_);
}

// calcDihedralAndSlopes, fb.cpp:2341-2453.  The angle and its slopes are
// produced in two calls so that the 5x3 slope table never has to be held:
// dihedral(..., 0, cs, sn) returns cos and sin of the angle (X/|(X,Y)|,
// Y/|(X,Y)|: the angle is atan2(Y, X)); dihedral(..., fscale, ...)
// recomputes the geometry and adds fscale * (y_slope + x_slope) of each of
// the five atoms CA(i-1), CA(i), CA(i+1), O(i-1), O(i) into the residue's
// local-force sums F (compute_rama_force, fb.cpp:2455-2506, summed over the
// wells).
template <class T>
AW_DEV void dihedral(const View<T> &V, int i, bool phi, double fscale, double &cs, double &sn, LocF &F)
{
  ParamPtr P = V.P;
  double an = P->an, bn = P->bn, cn = P->cn, ap = P->ap, bp = P->bp, cp = P->cp;
  V3<double> a, b, c;
  if (phi) {
    a = dvec<double>(XV(V, K_CP, i), XA(V, K_CA, i));
    b = dvec<double>(XA(V, K_CA, i), XV(V, K_N, i));
    c = dvec<double>(XV(V, K_N, i), XV(V, K_CP, i - 1));
  } else {
    a = dvec<double>(XV(V, K_N, i + 1), XV(V, K_CP, i));
    b = dvec<double>(XV(V, K_CP, i), XA(V, K_CA, i));
    c = dvec<double>(XA(V, K_CA, i), XV(V, K_N, i));
  }
  double adb = dot(a, b), bdc = dot(b, c), adc = dot(a, c), b2 = dot(b, b), bm = sqrt(b2);
  V3<double> bxa = crs(b, a);
  double cdbxa = dot(c, bxa);
  double Y = bm * cdbxa, X = adb * bdc - b2 * adc;
  if (fscale == 0.0) {
    double h = sqrt(X * X + Y * Y);
    if (h > 0.0) { cs = X / h; sn = Y / h; }
    else { cs = 1.0; sn = 0.0; }  // atan2(0, 0) = 0
    return;
  }
  V3<double> cxa = crs(c, a), cxb = crs(c, b);
  double X2Y2 = 1.0 / (X * X + Y * Y);
  double dy = X * X2Y2, dx = -Y * X2Y2;
  b2 *= dx; adb *= dx; adc *= dx; bdc *= dx;
  cdbxa *= dy / bm;
  bm *= dy;
#pragma unroll
  for (int l = 0; l < 3; l++) {
    double av = l == 0 ? a.x : (l == 1 ? a.y : a.z), bv = l == 0 ? b.x : (l == 1 ? b.y : b.z);
    double cv = l == 0 ? c.x : (l == 1 ? c.y : c.z);
    double bxav = l == 0 ? bxa.x : (l == 1 ? bxa.y : bxa.z), cxav = l == 0 ? cxa.x : (l == 1 ? cxa.y : cxa.z);
    double cxbv = l == 0 ? cxb.x : (l == 1 ? cxb.y : cxb.z);
    double h1, h2, h3, h4;
    double g1 = bv * bdc - cv * b2;
    double g2 = av * bdc - 2.0 * bv * adc + cv * adb;
    double g3 = bv * adb - av * b2;
    double o0, o1, o2, o3, o4;
    if (phi) {
      h1 = cxbv * bm; h2 = cxav * bm; h3 = bxav * bm; h4 = bv * cdbxa;
      o0 = -an * h4 + (an - ap) * h3 + an * h2 + (-an * g2 + (an - ap) * g3);
      o1 = (1.0 - bn) * h4 + (bn - bp) * h3 + (ap - 1.0) * h1 - (1.0 - bn) * h2 +
           ((ap - 1.0) * g1 + (1.0 - bn) * g2 + (bn - bp) * g3);
      o2 = bp * h1 + bp * g1;
      o3 = -cn * h4 + (cn - cp) * h3 + cn * h2 + (-cn * g2 + (cn - cp) * g3);
      o4 = cp * h1 + cp * g1;
    } else {
      h1 = bxav * bm; h2 = cxbv * bm; h3 = cxav * bm; h4 = bv * cdbxa;
      o0 = -an * h1 + (-an * g3);
      o1 = (ap - 1.0) * h4 + (1.0 - bn) * h1 + (an - ap) * h2 - (ap - 1.0) * h3 +
           ((an - ap) * g1 + (ap - 1.0) * g2 + (1.0 - bn) * g3);
      o2 = bp * h4 + (bn - bp) * h2 - bp * h3 + ((bn - bp) * g1 + bp * g2);
      o3 = -cn * h1 + (-cn * g3);
      o4 = cp * h4 + (cn - cp) * h2 - cp * h3 + ((cn - cp) * g1 + cp * g2);
    }
    F.f[LocF::slot(K_CA, -1)][l] += fscale * o0;
    F.f[LocF::slot(K_CA, 0)][l] += fscale * o1;
    F.f[LocF::slot(K_CA, 1)][l] += fscale * o2;
    F.f[LocF::slot(K_O, -1)][l] += fscale * o3;
    F.f[LocF::slot(K_O, 0)][l] += fscale * o4;
  }
}

// compute_rama_potential / compute_rama_force, fb.cpp:2455-2546
template <class T>
AW_DEV void rama_term(const View<T> &V, int i, int ri, EAcc &E, LocF &F)
{
  ParamPtr P = V.P;
  double cphi, sphi, cpsi, spsi, dummy;
  dihedral(V, i, true, 0.0, cphi, sphi, F);
  dihedral(V, i, false, 0.0, cpsi, spsi, F);
  // wells j = js + k, k < nk: the general set (js = 0) or, for a proline
  // with Rama_P on, the proline set (js = 6); at most 6 wells per set.  The
  // two sets run as two loops over a wave-uniform index, each with its
  // lanes only: every parameter is a scalar load, with no per-lane choice
  // between the sets.
  const bool pro = (ri & RI_PRO) && P->rama_p_flag;
  const int nr = P->n_rama_par, npr = P->n_rama_p_par;
  const AW_G double *apsr = V.aps + i * 12 + (pro ? 6 : 0);
  // the residue's six weights of its set (in its 12-entry row), read
  // together before the wells
  double apsk[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) apsk[k] = apsr[k];
  double fphi = 0, fpsi = 0;
  double ev = 0.0;
  auto well = [&](int j, double apsv) {
    const int ssw = P->ssweight[j];
    if (ssw && apsv == 0.0) return;
    const double cf = P->cphi0[j], sf = P->sphi0[j], cs = P->cpsi0[j], ss = P->spsi0[j];
    // cos/sin(phi + phi0) by angle addition
    double cp_ = cphi * cf - sphi * sf, sp_ = sphi * cf + cphi * sf;
    double cq_ = cpsi * cs - spsi * ss, sq_ = spsi * cs + cpsi * ss;
    double cph = cp_ - 1.0, cps = cq_ - 1.0;
    double pc = P->phiw[j] * cph, sc = P->psiw[j] * cps;
    double Vv = P->w[j] * exp(-cph * pc - cps * sc);
    if (ssw) Vv *= apsv;
    double force = 2.0 * Vv;
    fphi += force * pc * sp_;
    fpsi += force * sc * sq_;
    ev += -Vv;
  };
  if (!pro) {
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (k < nr) well(k, apsk[k]);
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k)
      if (k < npr) well(6 + k, apsk[k]);
  }
  E.rama += ev;
  if (fphi != 0.0) dihedral(V, i, true, fphi, dummy, dummy, F);
  if (fpsi != 0.0) dihedral(V, i, false, fpsi, dummy, dummy, F);
}

// LAMMPS bond_style harmonic on the deck's bonds (outside the reference).
// Always FP64: r - r0 cancels for bonds near their rest length.
template <class T>
AW_DEV void bond_one(const View<T> &V, int i, int ka, int da, int kb, int db, double r0, EAcc &E, LocF &F)
{
  V3<double> d = dvec<double>(XA(V, ka, i + da), XA(V, kb, i + db));
  double r = sqrt(dot(d, d)), dr = r - r0;
  double k = V.P->bond_k;
  E.bond += k * dr * dr;
  double fb = r > 0.0 ? -2.0 * k * dr / r : 0.0;
  F.add(ka, da, fb, d);
  F.add(kb, db, -fb, d);
}

template <class T>
AW_DEV void bonds_term(const View<T> &V, int i, int ri, EAcc &E, LocF &F)
{
  ParamPtr P = V.P;
  bond_one(V, i, K_CA, 0, K_O, 0, P->bond_r0[1], E, F);
  bond_one(V, i, K_CA, 0, K_CB, 0, (ri & RI_GLY) ? P->bond_r0[4] : P->bond_r0[3], E, F);
  if (!(ri & RI_LAST)) {
    bond_one(V, i, K_CA, 0, K_CA, 1, P->bond_r0[0], E, F);
    bond_one(V, i, K_O, 0, K_CA, 1, P->bond_r0[2], E, F);
  }
}

// ------------------------------------------------------------------------
// fragment memory: one item = one (i, j) pair of one memory group, looping
// over the group's memories (fb.cpp:4079-4177, Rf fm.cpp:190-202).  Pairs
// are grouped by window because all memories of a window share the target
// distances r_ij; the Rf of memory m sits at rf0 + m * stride.

template <class T> struct T4 {};
template <> struct T4<float> { typedef float4 type; typedef FMItem item; };
template <> struct T4<double> { typedef double4 type; typedef FMItemD item; };

// a 4-vector load from device memory as a global (not flat) access: the
// fragment-memory tables never live in LDS
template <class T>
AW_DEV typename T4<T>::type ld_global4(const typename T4<T>::type *p)
{
#if defined(__HIP_DEVICE_COMPILE__)
  typedef T V4 __attribute__((ext_vector_type(4)));
  V4 v = *(const __attribute__((address_space(1))) V4 *)p;
  return typename T4<T>::type(v.x, v.y, v.z, v.w);
#else
  return *p;
#endif
}

// one memory's four Gaussian wells of one pair: rf = (CA-CA, CA(i)-CB(j),
// CA(j)-CB(i), CB-CB) pre-scaled by the item's sc, AW_FM_FAR where the
// Gaussian is skipped
// EF: accumulate the energy too (only the evaluation whose energies are
// reported needs it)
// LW: w is log2 of the weight, which then enters the exponent
template <bool EF, bool LW, class T4t, class T>
AW_DEV void fm_gauss(const T4t &f, T w, const T *r, T &S0, T &S1, T &S2, T &S3, T &Es)
{
  // r and f are pre-scaled by the item's sc: exp2(-d^2) is the Gaussian
  T d0 = r[0] - f.x, d1 = r[1] - f.y, d2 = r[2] - f.z, d3 = r[3] - f.w;
  T v0 = LW ? t_exp2(w - d0 * d0) : w * t_exp2(-(d0 * d0));
  T v1 = LW ? t_exp2(w - d1 * d1) : w * t_exp2(-(d1 * d1));
  T v2 = LW ? t_exp2(w - d2 * d2) : w * t_exp2(-(d2 * d2));
  T v3 = LW ? t_exp2(w - d3 * d3) : w * t_exp2(-(d3 * d3));
  S0 += v0 * d0; S1 += v1 * d1; S2 += v2 * d2; S3 += v3 * d3;
  if (EF) Es += (v0 + v1) + (v2 + v3);
}

// force sinks of fm_run.  fm_run is compiled as a function of its own, so
// the sinks state where their arrays live (a base typed by aw_cast, or
// aadd_as) for the compiler to emit LDS or global atomics.
// FmSinkForce: adds into a replica's force array f[3][n][3]
template <bool FLDS>
struct FmSinkForce {
  typename aw_as<FLDS>::type *f;  // typed once (aw_cast): 32-bit LDS offsets when FLDS
  int n;
  template <class R>
  AW_DEV void operator()(int kind, int i, R fx, R fy, R fz) const
  {
    auto *p = f + (kind * n + i) * 3;
    aadd_t(p + 0, (double)fx);
    aadd_t(p + 1, (double)fy);
    aadd_t(p + 2, (double)fz);
  }
};

// a lane's fragment-memory items k, k + ks, ... (< k1), software-pipelined
// so that a lane keeps its table loads in flight while it computes:
//   - a descriptor is loaded an item ahead;
//   - the first batch (FMB memories) of an item is loaded while the previous
//     item computes its last batch, and an item's second batch (4 memories)
//     is loaded before its first batch computes.
// Memories beyond FMB + 4 (groups larger than the shipped 20 per window) are
// computed in further batches of 4 without overlap.  Loads past nmem re-read
// the last memory, whose weight there is the zero padding of coef (no
// branches on the data).  Coordinates x[3][n][3] (CA, CB, O) are in LDS when
// XLDS; forces go to sink(kind, residue, fx, fy, fz); returns the energy
// (EF; 0 otherwise).
// fm_run (below) is the same code as a function of its own (not inlined):
// it gets its own register allocation for the batches in flight, and is
// called once per pass; the wide engine inlines fm_run_body unpipelined.
template <class T>
struct FmBatch {
  static const int FMB = sizeof(T) == 4 ? 16 : 8;
};

// byte offsets of an item's table rows: memory m's Rf entry at ob + m * sb,
// its weights at cb (from the replica's table bases); nm1 = nmem - 1.  The
// tables of one replica stay below 4 GiB (checked when they are packed), so
// 32-bit offsets from a wave-uniform base address them.
struct FmRow {
  unsigned ob, sb, cb;
  int nm1;
};
template <class T>
AW_DEV FmRow fm_row(const typename T4<T>::item &it)
{
  FmRow a;
  a.ob = (unsigned)it.rf0 * (unsigned)sizeof(typename T4<T>::type);
  a.sb = (unsigned)it.stride * (unsigned)sizeof(typename T4<T>::type);
  a.cb = (unsigned)it.coef0 * (unsigned)sizeof(T);
  a.nm1 = (it.nmem & AW_FM_NMEM_MASK) - 1;
  return a;
}

// memories m0 .. m0 + nb - 1 of a row (nb a multiple of 4): weights from the
// zero-padded run of the group, Rf entries with m clamped to nmem - 1
template <class T>
AW_DEV void fm_load(const char *rfb, const char *cfb, const FmRow &a, int m0, int nb, typename T4<T>::type *f4,
                    T *cf)
{
  typedef typename T4<T>::type T4t;
  const T4t *cf4 = (const T4t *)(cfb + a.cb) + (m0 >> 2);
#pragma unroll
  for (int u = 0; u < nb / 4; ++u) {
    T4t c = ld_global4<T>(cf4 + u);
    cf[4 * u] = c.x; cf[4 * u + 1] = c.y; cf[4 * u + 2] = c.z; cf[4 * u + 3] = c.w;
  }
#pragma unroll
  for (int u = 0; u < nb; ++u) {
    int m = m0 + u < a.nm1 ? m0 + u : a.nm1;
    f4[u] = ld_global4<T>((const T4t *)(rfb + (a.ob + aw_umul24((unsigned)m, a.sb))));
  }
}

// PIPE = false (lanes with one or two items, as in the grid-wide phase
// engine): no prefetch of the next item's first batch, each item loads its
// own
// off: residue index of x's first row (items carry replica residue indices;
// a copy of a residue span is addressed as rows i - off)
template <class T, bool XLDS, bool EF, class Sink, bool PIPE = true, int FMB = FmBatch<T>::FMB, int TB = 4>
AW_DEV double fm_run_body(const double *x, int n, const void *rf, const T *coef,
                          const typename T4<T>::item *items, int k, int k1, int ks, Sink sink, int off = 0)
{
  static_assert(FMB % 4 == 0 && FMB <= AW_FM_PAD, "first batch: whole weight vectors within the padding");
  static_assert(TB % 4 == 0 && TB <= AW_FM_PAD, "second batch: whole weight vectors");
  typedef typename T4<T>::type T4t;
  typedef typename T4<T>::item ItemT;
  if (k >= k1) return 0.0;
  // the bases are the same in every lane: scalar registers
  const char *rfb = (const char *)aw_uniform(rf), *cfb = (const char *)aw_uniform(coef);
  x = aw_uniform(x);
  const auto *xs = aw_cast<XLDS>(x);  // coordinates in their own address space
  items = aw_uniform(items);
  double efm = 0.0;
  ItemT cur = ld_item(items + k);
  ItemT nxt = k + ks < k1 ? ld_item(items + k + ks) : cur;
  T4t fa[FMB];
  T ca[FMB];
  if (PIPE) fm_load(rfb, cfb, fm_row<T>(cur), 0, FMB, fa, ca);
  for (; k < k1; k += ks) {
    if (!PIPE) fm_load(rfb, cfb, fm_row<T>(cur), 0, FMB, fa, ca);
    int nmem = cur.nmem & AW_FM_NMEM_MASK;
    // second batch: memories FMB..FMB+TB-1, or (no such memories) a re-read
    // of the first TB with weight 0; loads and compute are unconditional so
    // that the wait for this batch leaves the next item's loads in flight
    T4t fb[TB];
    T cb[TB];
    bool tail = nmem > FMB;
    FmRow row = fm_row<T>(cur);
    fm_load(rfb, cfb, row, tail ? FMB : 0, TB, fb, cb);
    AW_SCHED_FENCE();  // the loads stay ahead of the first batch's arithmetic
    int i = (cur.ij >> 16) - off, j = (cur.ij & 0xffff) - off;
    bool gi = cur.nmem & AW_FM_GLY_I, gj = cur.nmem & AW_FM_GLY_J;
    double cai[3], caj[3], cbi[3], cbj[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      cai[c] = xs[3 * i + c];
      caj[c] = xs[3 * j + c];
      cbi[c] = xs[3 * (n + i) + c];
      cbj[c] = xs[3 * (n + j) + c];
    }
    V3<T> d[4];
    d[0] = dvec<T>(cai, caj);
    d[1] = dvec<T>(cai, cbj);
    d[2] = dvec<T>(cbi, caj);
    d[3] = dvec<T>(cbi, cbj);
    T r[4], rs[4];
    T sc = (T)cur.sc;
    for (int q = 0; q < 4; ++q) {
      r[q] = t_sqrt(dot(d[q], d[q]));
      rs[q] = r[q] * sc;
    }
    T S0 = 0, S1 = 0, S2 = 0, S3 = 0, Es = 0;
    bool lw = cur.nmem & AW_FM_LOGW;
    if (lw) {
#pragma unroll
      for (int u = 0; u < FMB; ++u) {
        fm_gauss<EF, true>(fa[u], ca[u], rs, S0, S1, S2, S3, Es);
        if ((u & 3) == 3) AW_SCHED_FENCE();  // bounds the interleaving (register pressure)
      }
    } else {
#pragma unroll
      for (int u = 0; u < FMB; ++u) {
        fm_gauss<EF, false>(fa[u], ca[u], rs, S0, S1, S2, S3, Es);
        if ((u & 3) == 3) AW_SCHED_FENCE();
      }
    }
    // the next item's first batch is in flight while this one finishes (past
    // the lane's last item, nxt is a valid item loaded again and unused)
    if (PIPE) fm_load(rfb, cfb, fm_row<T>(nxt), 0, FMB, fa, ca);
    // a second batch past the group's end weighs 0 (exp2(-inf) = 0 for log
    // weights)
    // (when the batch may reach past a group's zero-padded weight block, its
    // memories past nmem are masked explicitly)
    T w0 = lw ? -(T)INFINITY : (T)0;
    T wb[TB];
#pragma unroll
    for (int u = 0; u < TB; ++u)
      wb[u] = (tail && (FMB + TB <= AW_FM_PAD || FMB + u < nmem)) ? cb[u] : w0;
    if (lw) {
#pragma unroll
      for (int u = 0; u < TB; ++u) fm_gauss<EF, true>(fb[u], wb[u], rs, S0, S1, S2, S3, Es);
    } else {
#pragma unroll
      for (int u = 0; u < TB; ++u) fm_gauss<EF, false>(fb[u], wb[u], rs, S0, S1, S2, S3, Es);
    }
    for (int m0 = FMB + TB; m0 < nmem; m0 += 4) {
      fm_load(rfb, cfb, row, m0, 4, fb, cb);
      if (lw) {
#pragma unroll
        for (int u = 0; u < 4; ++u) fm_gauss<EF, true>(fb[u], cb[u], rs, S0, S1, S2, S3, Es);
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) fm_gauss<EF, false>(fb[u], cb[u], rs, S0, S1, S2, S3, Es);
      }
    }
    // the weights entered with the sign of cf; the energy is -cf e
    bool act1 = !gj, act2 = !gi, act3 = !gi && !gj;
    T g = -(T)cur.gamma, invs = (T)cur.invs;
    if (EF) efm += (double)(g * Es);
    T f0 = t_div(g * S0 * invs, r[0]);
    T f1 = act1 ? t_div(g * S1 * invs, r[1]) : (T)0;
    T f2 = act2 ? t_div(g * S2 * invs, r[2]) : (T)0;
    T f3 = act3 ? t_div(g * S3 * invs, r[3]) : (T)0;
    sink(K_CA, i, f0 * d[0].x + f1 * d[1].x, f0 * d[0].y + f1 * d[1].y, f0 * d[0].z + f1 * d[1].z);
    sink(K_CA, j, -(f0 * d[0].x + f2 * d[2].x), -(f0 * d[0].y + f2 * d[2].y), -(f0 * d[0].z + f2 * d[2].z));
    if (act2) sink(K_CB, i, f2 * d[2].x + f3 * d[3].x, f2 * d[2].y + f3 * d[3].y, f2 * d[2].z + f3 * d[3].z);
    if (act1) sink(K_CB, j, -(f1 * d[1].x + f3 * d[3].x), -(f1 * d[1].y + f3 * d[3].y), -(f1 * d[1].z + f3 * d[3].z));
    cur = nxt;
    if (k + 2 * ks < k1) nxt = ld_item(items + k + 2 * ks);
  }
  return efm;
}

// the same as a function of its own (the resident engine's call: its own
// register allocation for the batches in flight)
template <class T, bool XLDS, bool EF, class Sink, bool PIPE = true>
AW_NOINLINE double fm_run(const double *x, int n, const void *rf, const T *coef,
                          const typename T4<T>::item *items, int k, int k1, int ks, Sink sink, int off = 0)
{
  // FP32: a first batch of 12 and a second of 8 (the shipped 20 memories
  // per window): the second batch's arithmetic covers more of the next
  // item's load latency
  const int FMB = sizeof(T) == 4 ? 12 : FmBatch<T>::FMB, TB = sizeof(T) == 4 ? 8 : 4;
  return fm_run_body<T, XLDS, EF, Sink, PIPE, FMB, TB>(x, n, rf, coef, items, k, k1, ks, sink, off);
}

#ifndef AW_HOST_EMU
// a shared replica's own workgroup: each wave claims chunks one at a time
// (the claim word's low half is the next unclaimed chunk) and evaluates a
// claimed chunk's items, a lane per item, into the replica's forces; the
// claimed chunks are marked in the workgroup's LDS mask.  Returns the
// energy of the lane's items (ef).
template <class T, bool XLDS>
AW_DEV double fm_claims(const View<T> &V, const typename T4<T>::item *items, bool ef)
{
  typedef FmSinkForce<XLDS> Sk;
  int lane = threadIdx.x & 63;
  double e = 0.0;
  auto claim = [&]() {
    unsigned c = 0;
    if (lane == 0)
      c = (unsigned)__hip_atomic_fetch_add(V.sh_word, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return c;
  };
  unsigned cn = claim();
  for (;;) {
    unsigned c = __builtin_amdgcn_readfirstlane(cn);
    if (c >= (unsigned)V.sh_n) break;
    cn = claim();  // the next claim is in flight while this chunk computes
    if (lane == 0)
      __hip_atomic_fetch_or(V.sh_mask, 1ull << c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    int k0 = __builtin_amdgcn_readfirstlane(V.sh_chunks[c].k0);
    int k1 = __builtin_amdgcn_readfirstlane(V.sh_chunks[c].k1);
    Sk sink{aw_cast<XLDS>(V.f()), V.n};
    if (ef)
      e += fm_run<T, XLDS, true, Sk, true>(V.x(), V.n, V.rf, V.coef, items, k0 + lane, k1, 64, sink);
    else
      fm_run<T, XLDS, false, Sk, true>(V.x(), V.n, V.rf, V.coef, items, k0 + lane, k1, 64, sink);
  }
  return e;
}
#else
template <class T, bool XLDS>
inline double fm_claims(const View<T> &, const void *, bool) { return 0.0; }
#endif

// ------------------------------------------------------------------------
// switching functions

// cWell::compute_theta_pair, sml.h:594-610 (theta and theta'/r).  r is
// passed in FP64 so that the small arguments kappa*(r - r_min) and
// kappa*(r_max - r) are formed before any rounding to FP32.
template <class T>
AW_DEV void well_theta(double kappa, double rmin, double rmax, double r, T &th, T &prd);

template <>
AW_DEV void well_theta<double>(double kappa, double rmin, double rmax, double r, double &th, double &prd)
{
  double tmin = tanh(kappa * (r - rmin)), tmax = tanh(kappa * (rmax - r));
  th = 0.25 * (1.0 + tmin) * (1.0 + tmax);
  prd = kappa * th * (tmax - tmin) / r;
}
// FP32: the same function through exponentials, without the cancellation
// of 1 + tanh(x) for x << 0: theta = 1/((1+e^-2a)(1+e^-2b)),
// theta'/r = 2 kappa (e^-2a - e^-2b) / (D^2 r)
template <>
AW_DEV void well_theta<float>(double kappa, double rmin, double rmax, double r, float &th, float &prd)
{
  float ea = aw_fast_exp((float)(-2.0 * kappa * (r - rmin))), eb = aw_fast_exp((float)(-2.0 * kappa * (rmax - r)));
  float D = (1.0f + ea) * (1.0f + eb);
  float iD = aw_frcp(D);
  th = iD;
  prd = t_div(2.0f * (float)kappa * (ea - eb) * iD * iD, (float)r);
}

template <class T>
AW_DEV void wtheta(ParamPtr P, int w, double r, T &th, T &prd)
{
  well_theta<T>(P->water_kappa, P->well_r_min[w], P->well_r_max[w], r, th, prd);
}
template <class T>
AW_DEV void htheta(ParamPtr P, double r, T &th, T &prd)
{
  well_theta<T>(P->helix_kappa, P->helix_r_min, P->helix_r_max, r, th, prd);
}

AW_DEV const double *site(const View<float> &V, int i, int ri) { return (ri & RI_GLY) ? XA(V, K_CA, i) : XA(V, K_CB, i); }
AW_DEV const double *site(const View<double> &V, int i, int ri) { return (ri & RI_GLY) ? XA(V, K_CA, i) : XA(V, K_CB, i); }

// cP_AP::compute (sml.h:176-203) in FP64: the nu < pap_delta early exit of
// fb.cpp:3281 is a hard threshold
// nu = (1 + tanh a) / 2 and prd = pref (1 - tanh^2 a) / (2 dr), a formed in
// FP64 (the P_AP and DSSP switches).  FP32: through E = exp(-2|a|), nu = 1/(1+E) (a >= 0) or E/(1+E),
// 1 - tanh^2 = 4E/(1+E)^2, without the cancellation of 1 + tanh for a << 0
template <class T>
AW_DEV void tanh_switch(double a, double pref, double dr, T &nu, T &prd);
template <>
AW_DEV void tanh_switch<double>(double a, double pref, double dr, double &nu, double &prd)
{
  double th = tanh(a);
  nu = 0.5 * (1 + th);
  prd = 0.5 * pref * (1 - th * th) / dr;
}
template <>
AW_DEV void tanh_switch<float>(double a, double pref, double dr, float &nu, float &prd)
{
  float af = (float)a;
  float e = aw_fast_exp(-2.0f * (af < 0.0f ? -af : af));
  float ip = aw_frcp(1.0f + e);
  nu = af >= 0.0f ? ip : e * ip;
  prd = 2.0f * (float)pref * e * ip * ip * aw_frcp((float)dr);
}

// the P_AP switch's parameters, read once by the caller
struct PapP {
  double drmax2, drmin2, pref, cut;
};
AW_DEV PapP pap_p(ParamPtr P)
{
  PapP q;
  q.drmax2 = P->pap_drmax_sq; q.drmin2 = P->pap_drmin_sq; q.pref = P->pap_pref; q.cut = P->pap_cut;
  return q;
}
template <class T>
AW_DEV void pap_nu(const PapP &q, const double *a, const double *b, T &nu, T &prd)
{
  double drsq = rsq_exact(a, b);
  if (drsq > q.drmax2) { nu = 0; prd = 0; }
  else if (drsq < q.drmin2) { nu = 1; prd = 0; }
  else {
    double dr = sqrt_t<T>(drsq);
    tanh_switch<T>(q.pref * (q.cut - dr), q.pref, dr, nu, prd);
  }
}

// the nu < 1e-12 early exit of compute_P_AP_potential: an FP32 nu decides
// unless it lies within a relative 1e-5 of the threshold, where FP64 decides
template <class T>
AW_DEV bool pap_keep(const PapP &q, const double *a, const double *b, T nu)
{
  if (sizeof(T) == 8) return !(nu < (T)AW_PAP_DELTA);
  double d = (double)nu - AW_PAP_DELTA;
  if (d > 1e-5 * AW_PAP_DELTA) return true;
  if (d < -1e-5 * AW_PAP_DELTA) return false;
  double nud, prdd;
  double drsq = rsq_exact(a, b);
  double dr = sqrt(drsq);
  tanh_switch<double>(q.pref * (q.cut - dr), q.pref, dr, nud, prdd);
  return !(nud < AW_PAP_DELTA);
}

// DSSP nu switch (fb.cpp:3006-3045): nu = (1 + tanh(pref (r - d0))) / 2,
// prd = pref nu (1 - tanh) / r; FP64 in the reference's form, FP32 through
// tanh_switch
template <class T>
AW_DEV void dssp_nu(double pref, double r, double d0, T &nu, T &prd)
{
  if (sizeof(T) == 8) {
    double th = tanh(pref * (r - d0));
    double v = 0.5 * (1.0 + th);
    nu = (T)v;
    prd = (T)(pref * v * (1.0 - th) / r);
  } else {
    tanh_switch<T>(pref * (r - d0), pref, r, nu, prd);
  }
}
// nu < 1e-4, decided in FP64 where an FP32 nu lies within a relative 1e-5
// of the threshold (nu = 0 and 1 are exact)
template <class T>
AW_DEV bool dssp_nu_below(T nu, double pref, double r, double d0)
{
  if (sizeof(T) == 8) return nu < (T)AW_DSSP_NU_DELTA;
  double d = (double)nu - AW_DSSP_NU_DELTA;
  if (d > 1e-5 * AW_DSSP_NU_DELTA) return false;
  if (d < -1e-5 * AW_DSSP_NU_DELTA) return true;
  return 0.5 * (1.0 + tanh(pref * (r - d0))) < AW_DSSP_NU_DELTA;
}

// ------------------------------------------------------------------------
// DSSP hydrogen bond (acceptor O of i, donor N-H of j), fb.cpp:2902-3277

template <class T>
AW_DEV void dssp_term(const View<T> &V, int i, int j, EAcc &E)
{
  ParamPtr P = V.P;
  // the scalar parameters, read before any branch (invariant loads that
  // leave the queue loop)
  const double kd = P->k_dssp, NOz = P->NO_zero, HOz = P->HO_zero, sNO = P->sigma_NO_sqinv, sHO = P->sigma_HO_sqinv;
  const double pr0 = P->pref[0], pr1 = P->pref[1], dnu0 = P->d_nu0, cut1 = P->nu_cut1_sq, cut2 = P->nu_cut2_sq;
  const double an_ = P->an, bn_ = P->bn, cn_ = P->cn, ah_ = P->ah, bh_ = P->bh, ch_ = P->ch;
  const int nr = P->n_rama_par;
  const int *RI = V.rinfo;
  const int n = V.n;
  int ri = RI[i], rj = RI[j];
  // neighbours' residue info; indices clamped into the system (a clamped
  // entry is never used: i is not a chain's last residue, j not its first)
  int rip1 = RI[i + 1 < n ? i + 1 : i], rim1 = RI[i > 0 ? i - 1 : i];
  int rjp1 = RI[j + 1 < n ? j + 1 : j], rjm1 = RI[j > 0 ? j - 1 : j];
  int rip2 = RI[i + 2 < n ? i + 2 : i], rim2 = RI[i >= 2 ? i - 2 : i];
  int rjp2 = RI[j + 2 < n ? j + 2 : j], rjm2 = RI[j >= 2 ? j - 2 : j];
  int ic = rchain(ri), jc = rchain(rj);
  // chains are contiguous index ranges: i +- 2 lie in i's chain when the
  // residues there carry its chain id (no chain-bound table read)
  bool ip2_in = i + 2 < n && rchain(rip2) == ic, im2_in = i >= 2 && rchain(rim2) == ic;
  bool jp2_in = j + 2 < n && rchain(rjp2) == jc, jm2_in = j >= 2 && rchain(rjm2) == jc;
  int ti = rtype(ri), tj = rtype(rj), tip1 = rtype(rip1), tim1 = rtype(rim1), tjp1 = rtype(rjp1), tjm1 = rtype(rjm1);
  int dij = j - i < 0 ? i - j : j - i;
  // sequence-separation class (fb.cpp:2956-3002)
  int cls = (ic == jc && dij < 45) ? (dij < 4 ? 0 : (dij < 18 ? 1 : 2)) : 3;
  int hb_class = cls + 1;
  // every table entry the pair can use, read before any branch: the
  // per-lane gathers then share one memory round trip instead of waiting
  // one after another
  const double aHB0 = P->anti_HB[ti][tj][0], aHB1 = P->anti_HB[ti][tj][1];
  const double aNHB0 = P->anti_NHB[tip1][tjm1][0] + P->anti_NHB[tim1][tjp1][0];
  const double aNHB1 = P->anti_NHB[tip1][tjm1][1] + P->anti_NHB[tim1][tjp1][1];
  const double pHB0 = P->para_HB[tip1][tj][0], pHB1 = P->para_HB[tip1][tj][1];
  const double aone = P->anti_one[ti] + P->anti_one[tj], pone = P->para_one[tip1] + P->para_one[tj];
  const int nrk = nr > 0 ? nr - 1 : 0;
  const double apsi = V.aps[i * 12 + nrk], apsj = V.aps[j * 12 + nrk];
  auto hr = P->hbscl[cls];
  double h[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) h[k] = hr[k];
  bool jlast = rj & RI_LAST;
  bool i_rep = !(jlast || (rjp1 & RI_PRO));
  bool i_AP = !((ri & RI_FIRST) || jlast || (ri & RI_PRO));
  // i >= chain end - 2 (fb.cpp) is i + 2 outside the chain
  bool i_P = !(!ip2_in || jlast || (rip2 & RI_PRO));
  T tsa_HB[2] = {0, 0}, tsa_NHB[2] = {0, 0}, tsp_HB[2] = {0, 0};
  if (i_AP) {
    tsa_HB[0] = (T)(0.5 * aHB0); tsa_HB[1] = (T)(0.5 * aHB1);
    tsa_NHB[0] = (T)(0.25 * aNHB0); tsa_NHB[1] = (T)(0.25 * aNHB1);
  }
  if (i_P) { tsp_HB[0] = (T)pHB0; tsp_HB[1] = (T)pHB1; }
  T lambda[4] = {0, 0, 0, 0};
  if (cls == 0) {
    lambda[0] = (T)-h[0]; lambda[1] = (T)-h[1];
  } else if (cls == 1) {
    if (nr > 0 && apsi == 1.0 && apsj == 1.0) {
      lambda[0] = (T)-h[0];
      lambda[1] = (T)-h[1];
      lambda[2] = (T)-h[2] - (T)h[3] * tsa_HB[0] - (T)h[4] * tsa_NHB[0] - (T)(h[5] * aone);
      lambda[3] = (T)-h[6];
    } else {
      lambda[1] = (T)-h[1];
    }
  } else {
    lambda[0] = (T)-h[0];
    lambda[1] = (T)-h[1];
    lambda[2] = (T)-h[2] - (T)h[3] * tsa_HB[1] - (T)h[4] * tsa_NHB[1] - (T)(h[5] * aone);
    lambda[3] = (T)-h[6] - (T)h[7] * tsp_HB[1] - (T)(h[8] * pone);
  }
  // nu switches (arguments formed in FP64); the nu < 1e-4 early exit
  // (fb.cpp:3052) is a hard threshold, decided as in FP64 (dssp_nu_below)
  T nu0 = 0, prdnu0 = 0, nu1 = 0, prdnu1 = 0;
  double r0 = 0.0, r1 = 0.0;
  bool nu0_on = (im2_in && ip2_in && hb_class != 2);
  bool nu1_on = (jm2_in && jp2_in && hb_class != 2);
  if (nu0_on) {
    double rr = rsq_exact(XA(V, K_CA, i + 2), XA(V, K_CA, i - 2));
    if (rr > cut1) {
      r0 = sqrt(rr);
      dssp_nu<T>(pr0, r0, dnu0, nu0, prdnu0);
    }
  } else nu0 = 1;
  if (nu1_on) {
    double rr = rsq_exact(XA(V, K_CA, j + 2), XA(V, K_CA, j - 2));
    if (rr > cut2) {
      r1 = sqrt(rr);
      dssp_nu<T>(pr1, r1, dnu0, nu1, prdnu1);
    }
  } else nu1 = 1;
  if (dssp_nu_below(nu0, pr0, r0, dnu0) || dssp_nu_below(nu1, pr1, r1, dnu0)) return;

  bool on[4] = {true, i_rep, i_AP, i_P};
  int acc[4] = {i, i, j, j}, don[4] = {j, j + 1, i, i + 2};
  T theta[4], p0[4], p1[4];
  V3<T> uk[4], wk[4];
  for (int k = 0; k < 4; ++k) {
    theta[k] = 0; p0[k] = 0; p1[k] = 0;
    if (!on[k]) continue;
    // distances and their offsets from the well centres in FP64 (the
    // offsets are small differences of ~3 A quantities)
    const double *oa = XA(V, K_O, acc[k]), *nd = XV(V, K_N, don[k]), *hd = XV(V, K_H, don[k]);
    double RNO = sqrt_t<T>(rsq_exact(oa, nd));
    double RHO = sqrt_t<T>(rsq_exact(oa, hd));
    // the force directions, from the same reads (the force loop below
    // follows atomic adds into the slab, after which they would be read
    // and subtracted again)
    uk[k] = dvec<T>(oa, nd);
    wk[k] = dvec<T>(oa, hd);
    T eNO = (T)(RNO - NOz), eHO = (T)(RHO - HOz);
    T dNO = eNO * (T)sNO;
    T dHO = eHO * (T)sHO;
    theta[k] = t_exp((T)-0.5 * (eNO * dNO + eHO * dHO));
    p0[k] = t_div(-dNO, (T)RNO);
    p1[k] = t_div(-dHO, (T)RHO);
  }
  T ff1 = (T)kd * theta[0];
  T Vv[4] = {lambda[0], lambda[1] * theta[1], lambda[2] * theta[2], lambda[3] * theta[3]};
  T theta_sum = ff1 * (Vv[0] + Vv[1] + Vv[2] + Vv[3]);
  ff1 *= nu0 * nu1;
  for (int k = 0; k < 4; ++k) Vv[k] *= ff1;
  T VT = Vv[0] + Vv[1] + Vv[2] + Vv[3];
  E.dssp += (double)VT;
  if (nu0_on) {
    V3<T> dn = dvec<T>(XA(V, K_CA, i + 2), XA(V, K_CA, i - 2));
    T fq = theta_sum * (prdnu0 * nu1);
    fadd(V, K_CA, i - 2, fq, dn);
    fadd(V, K_CA, i + 2, -fq, dn);
  }
  if (nu1_on) {
    V3<T> dn = dvec<T>(XA(V, K_CA, j + 2), XA(V, K_CA, j - 2));
    T fq = theta_sum * (nu0 * prdnu1);
    fadd(V, K_CA, j - 2, fq, dn);
    fadd(V, K_CA, j + 2, -fq, dn);
  }
  Vv[0] = VT;
  T an = (T)an_, bn = (T)bn_, cn = (T)cn_, ah = (T)ah_, bh = (T)bh_, ch = (T)ch_;
  for (int k = 0; k < 4; ++k) {
    if (!on[k]) continue;
    T q0 = p0[k] * Vv[k], q1 = p1[k] * Vv[k];
    int a = acc[k], d = don[k];
    const V3<T> &u = uk[k], &w = wk[k];
    fadd3(V, K_O, a, -(q0 * u.x + q1 * w.x), -(q0 * u.y + q1 * w.y), -(q0 * u.z + q1 * w.z));
    fadd3(V, K_CA, d - 1, an * q0 * u.x + ah * q1 * w.x, an * q0 * u.y + ah * q1 * w.y, an * q0 * u.z + ah * q1 * w.z);
    fadd3(V, K_CA, d, bn * q0 * u.x + bh * q1 * w.x, bn * q0 * u.y + bh * q1 * w.y, bn * q0 * u.z + bh * q1 * w.z);
    fadd3(V, K_O, d - 1, cn * q0 * u.x + ch * q1 * w.x, cn * q0 * u.y + ch * q1 * w.y, cn * q0 * u.z + ch * q1 * w.z);
  }
}

// ------------------------------------------------------------------------
// P_AP (i < j by residue index), fb.cpp:3279-3383

template <class T>
AW_DEV void pap_term(const View<T> &V, int i, int j, EAcc &E)
{
  ParamPtr P = V.P;
  // parameters read before any branch (invariant loads leave the loop)
  const PapP q = pap_p(P);
  const int d = P->i_diff_pap, mn = P->i_med_min, mx = P->i_med_max, nr = P->n_rama_par;
  const double kp0 = P->k_pap[0], kp1 = P->k_pap[1], kp2 = P->k_pap[2];
  const double kbeta = P->k_betapred_pap, kglob = P->k_global_pap;
  // the pair's global reads (chain bounds, secondary-structure weights),
  // issued before the switch so that their latency overlaps it
  const int *RI = V.rinfo;
  int n = V.n;
  int ic = rchain(RI[i]), jc = rchain(RI[j]);
  const int ics = V.ch_start[ic], icl = V.ch_len[ic], jcs = V.ch_start[jc], jcl = V.ch_len[jc];
  const int nrk = nr > 0 ? nr - 1 : 0;
  const double apsi = V.aps[i * 12 + nrk], apsj = V.aps[j * 12 + nrk];
  // the CA positions the pair can use, read together before any branch
  // (indices clamped; i + d, j - d, j + d are used only where they exist):
  // read where used, they would wait behind the force adds before them
  auto clampi = [n](int k) { return k < 0 ? 0 : (k >= n ? n - 1 : k); };
  double ci[3], cj[3], cid[3], cjm[3], cjp[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ci[c] = XA(V, K_CA, i)[c]; cj[c] = XA(V, K_CA, j)[c];
    cid[c] = XA(V, K_CA, clampi(i + d))[c];
    cjm[c] = XA(V, K_CA, clampi(j - d))[c];
    cjp[c] = XA(V, K_CA, clampi(j + d))[c];
  }
  T nu_ij, prd_ij;
  pap_nu(q, ci, cj, nu_ij, prd_ij);
  if (!pap_keep(q, ci, cj, nu_ij)) return;
  int i_ch_end = ics + icl, j_ch_end = jcs + jcl;
  int jhi = (i + mx + 2 * d) < (i_ch_end - 1) ? (i + mx + 2 * d) : (i_ch_end - 1);
  bool med = ic == jc && i < i_ch_end - (mn + 2 * d) && j >= i + (mn + 2 * d) && j <= jhi;
  bool lng = (ic == jc && i < i_ch_end - (mx + 2 * d + 1) && j >= i + (mx + 2 * d + 1) && j < j_ch_end) ||
             (ic != jc && i + d < n && j - d >= 0 && rchain(RI[i + d]) == ic && rchain(RI[j - d]) == jc);
  bool par = (ic == jc && i < i_ch_end - (mx + 1 + d) && j >= i + (mx + 1) && j < i_ch_end - d) ||
             (ic != jc && i + d < n && j + d < n && rchain(RI[i + d]) == ic && rchain(RI[j + d]) == jc);
  bool beta = nr > 0 && apsi == 1.0 && apsj == 1.0;
  if (med || lng) {
    double K = beta ? (med ? kp0 : 0.0) + (lng ? kp1 * kbeta : 0.0)
                    : (med ? kp0 : 0.0) + (lng ? kp1 : 0.0);
    T nu2, prd2;
    pap_nu(q, cid, cjm, nu2, prd2);
    double kk = kglob * K;
    E.pap += -kk * (double)nu_ij * (double)nu2;
    V3<T> d0 = dvec<T>(ci, cj), d1 = dvec<T>(cid, cjm);
    T f0 = (T)kk * prd_ij * nu2, f1 = (T)kk * nu_ij * prd2;
    fadd(V, K_CA, i, -f0, d0); fadd(V, K_CA, j, f0, d0);
    fadd(V, K_CA, i + d, -f1, d1); fadd(V, K_CA, j - d, f1, d1);
  }
  if (par) {
    double K = beta ? kp2 * kbeta : kp2;
    T nu2, prd2;
    pap_nu(q, cid, cjp, nu2, prd2);
    double kk = kglob * K;
    E.pap += -kk * (double)nu_ij * (double)nu2;
    V3<T> d0 = dvec<T>(ci, cj), d1 = dvec<T>(cid, cjp);
    T f0 = (T)kk * prd_ij * nu2, f1 = (T)kk * nu_ij * prd2;
    fadd(V, K_CA, i, -f0, d0); fadd(V, K_CA, j, f0, d0);
    fadd(V, K_CA, i + d, -f1, d1); fadd(V, K_CA, j + d, f1, d1);
  }
}

// ------------------------------------------------------------------------
// excluded volume pair style, pev.cpp:50-138 (1-2 bonded pairs removed by
// special_bonds fene: CA(i)-CA(i+1), CA(i)-O(i), O(i)-CA(i+1), CA(i)-CB(i))

AW_DEV int atype(int kind, int ri) { return kind == K_CA ? 1 : (kind == K_O ? 3 : ((ri & RI_GLY) ? 5 : 4)); }
template <class TabPtr>
AW_DEV double ev_pick(TabPtr tab, int ta0, int ta1, int tb0, int tb1, bool gi, bool gj)
{
  double v00 = aw_sval(tab[ta0][tb0]), v01 = aw_sval(tab[ta0][tb1]), v10 = aw_sval(tab[ta1][tb0]), v11 = aw_sval(tab[ta1][tb1]);
  return gi ? (gj ? v11 : v10) : (gj ? v01 : v00);
}

template <class T>
AW_DEV void ev_pair(const View<T> &V, int i, int j, int ri, int rj, EAcc &E)
{
  ParamPtr P = V.P;
  const int p = P->ev_p;  // read before any branch (leaves the loop)
  bool same = rchain(ri) == rchain(rj);
  bool nb = same && (j == i + 1);
  bool shortc = same && (j - i < 5);
  bool gi = ri & RI_GLY, gj = rj & RI_GLY;
  // the six positions, read together before any test: read inside the
  // loop, each would wait behind the force adds of the combinations before
  // it (the compiler cannot move a read above a possibly aliasing add)
  double pa[3][3], pb[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) { pa[k][c] = XA(V, k, i)[c]; pb[k][c] = XA(V, k, j)[c]; }
#pragma unroll
  for (int ka = 0; ka < 3; ++ka)
#pragma unroll
    for (int kb = 0; kb < 3; ++kb) {
      if (nb && kb == K_CA && (ka == K_CA || ka == K_O)) continue;
      // parameters at wave-uniform indices (the glycine and non-glycine
      // entries, then a per-lane choice): scalar loads
      int ta0 = atype(ka, 0), ta1 = atype(ka, RI_GLY), tb0 = atype(kb, 0), tb1 = atype(kb, RI_GLY);
      double lam = ev_pick(P->ev_lambda, ta0, ta1, tb0, tb1, gi, gj);
      if (lam == 0.0) continue;
      double rc = shortc ? ev_pick(P->ev_cut_short, ta0, ta1, tb0, tb1, gi, gj)
                         : ev_pick(P->ev_cut_long, ta0, ta1, tb0, tb1, gi, gj);
      // FP32 estimate first: beyond rc^2 (1 + margin) the exact test fails too
      if (rsq_f(pa[ka], pb[kb]) > (float)(rc * rc) * (1.0f + AW_RSQ_MARGIN)) continue;
      double rsq = rsq_exact(pa[ka], pb[kb]);
      double r = sqrt(rsq);
      if (!(r < rc)) continue;
      V3<T> d = dvec<T>(pa[ka], pb[kb]);
      T rr = (T)r, dr = rr - (T)rc;
      T pw = (T)1;
      for (int q = 0; q < p - 1; ++q) pw *= dr;
      T pref = (T)((p % 2 == 0) ? lam : -lam) * pw;
      T fp = rr > (T)0 ? t_div(-pref * (T)p, rr) : (T)0;
      fadd(V, ka, i, fp, d);
      fadd(V, kb, j, -fp, d);
      E.ev += (double)(pref * dr);
    }
}

// same-residue excluded volume (only CB-O can be unbonded within a residue)
template <class T>
AW_DEV void ev_self(const View<T> &V, int i, int ri, EAcc &E, LocF &F)
{
  ParamPtr P = V.P;
  int ta = atype(K_CB, ri), tb = 3;
  double lam = P->ev_lambda[ta][tb];
  if (lam == 0.0) return;
  double rc = P->ev_cut_short[ta][tb];
  if (rsq_f(XA(V, K_CB, i), XA(V, K_O, i)) > (float)(rc * rc) * (1.0f + AW_RSQ_MARGIN)) return;
  double r = sqrt(rsq_exact(XA(V, K_CB, i), XA(V, K_O, i)));
  if (!(r < rc)) return;
  V3<T> d = dvec<T>(XA(V, K_CB, i), XA(V, K_O, i));
  T rr = (T)r, dr = rr - (T)rc;
  int p = P->ev_p;
  T pw = (T)1;
  for (int q = 0; q < p - 1; ++q) pw *= dr;
  T pref = (T)((p % 2 == 0) ? lam : -lam) * pw;
  T fp = rr > (T)0 ? t_div(-pref * (T)p, rr) : (T)0;
  F.add(K_CB, 0, fp, d);
  F.add(K_O, 0, -fp, d);
  E.ev += (double)(pref * dr);
}

// conservative pair prefilter on CA-CA distance: every atom of residue i
// lies within ext[i] of CA(i)
template <class T>
AW_DEV bool near_pair(const View<T> &V, int i, int j, float reach)
{
  const double *a = XA(V, K_CA, i), *b = XA(V, K_CA, j);
  float dx = (float)(a[0] - b[0]), dy = (float)(a[1] - b[1]), dz = (float)(a[2] - b[2]);
  float lim = reach + V.ext()[i] + V.ext()[j] + 1e-3f;
  return dx * dx + dy * dy + dz * dz <= lim * lim * 1.0001f;
}

// ------------------------------------------------------------------------
// Verlet list of residue pairs: every pair whose CA-CA distance is within
// pair_reach + ext_i + ext_j + skin at the build.  It stays a superset of
// the pairs any term can select while every residue's CA drift plus extent
// growth is below skin/2; the passes still apply each term's exact
// predicate.  Rebuilt on demand (phase_nl_build) by the resident engine.


// true when the list must be rebuilt (uniform; after phase A's barrier)
template <class T>
AW_DEV bool nl_needs_build(const View<T> &V)
{
  return V.nl && (V.nl_state[7] == 0 || i2f(V.nl_state[0]) > 0.5f * (float)V.P->nl_skin);
}

template <class T>
AW_DEV void phase_nl_build(const View<T> &V, int t0, int ts)
{
  int n = V.n, np = pair_count(n);
  float reach = (float)V.P->pair_reach + (float)V.P->nl_skin;
  for (int k = t0; k < np; k += ts) {
    int i, j;
    pair_decode(k, n, i, j);
    // the kept pairs of a wave reserve their slots with one atomic
    bool keep = near_pair(V, i, j, reach);
    lmask m = w_ballot(keep);
    if (m) {
      int base = q_reserve1(&V.nl_state[1], w_count(m));
      int q = base + w_rank(m);
      if (keep && q < V.nl_cap) V.nl[AW_NL_HDR + q] = (i << 16) | j;
    }
  }
  for (int i = t0; i < n; i += ts) {
    const double *ca = XA(V, K_CA, i);
    V.nlx[3 * i] = ca[0]; V.nlx[3 * i + 1] = ca[1]; V.nlx[3 * i + 2] = ca[2];
    V.nlx[3 * n + i] = (double)V.ext()[i];
  }
}

// launch start (one thread, after any reset of the header): the header's
// count and validity into nl_state[7]
template <class T>
AW_DEV void nl_state_init(const View<T> &V)
{
  V.nl_state[7] = V.nl[1] ? V.nl[0] + 1 : 0;
}

// after the build's barrier: publish count and validity (one thread)
template <class T>
AW_DEV void nl_publish(const View<T> &V, bool built)
{
  if (built) {
    int c = V.nl_state[1];
    V.nl[0] = c < V.nl_cap ? c : V.nl_cap;
    V.nl[1] = c <= V.nl_cap ? 1 : 0;
    V.nl[2] += 1;
    // the same count and validity for this launch's passes, next to the
    // queue counters (the header stays the host's record)
    V.nl_state[7] = c <= V.nl_cap ? c + 1 : 0;
  }
  V.nl_state[0] = 0;
  V.nl_state[1] = 0;
  V.nl_state[2] = V.nl_state[3] = V.nl_state[4] = V.nl_state[5] = V.nl_state[6] = 0;
}

// iterate the residue pairs of a pass: the Verlet list when valid, else all
// pairs; pairs [k0, k1) of that sequence (one call site of the body)
template <class T>
AW_DEV int pair_total(const View<T> &V)
{
  int c = V.nl ? V.nl_state[7] : 0;
  return c > 0 ? c - 1 : pair_count(V.n);
}

// body(i, j) for the packed pairs src[k0 + t0], src[k0 + t0 + ts], ...: the
// next entry is loaded while the current pair is evaluated, so a lane does
// not wait on a memory round trip per pair
template <class Src, class F>
AW_DEV void for_packed(Src src, int k0, int k1, int t0, int ts, F &&body)
{
  int k = k0 + t0;
  int cur = k < k1 ? src[k] : 0;
  for (; k < k1; k += ts) {
    // unconditional (clamped to the last entry): a branch around the load
    // would leave a path with no younger load, and the wait for cur would
    // then become a wait for every outstanding load at the top of the body
    int nxt = src[k + ts < k1 ? k + ts : k1 - 1];
    body(cur >> 16, cur & 0xffff);
    cur = nxt;
  }
}

// for_packed for a body that stores to global memory: body(i, j, nxt)
// gets the prefetched next entry and must aw_hold(nxt) on every path before
// its stores.  Vector-memory operations are counted in issue order, so a
// wait for nxt after this pair's stores would also wait for the stores.
template <class Src, class F>
AW_DEV void for_packed_st(Src src, int k0, int k1, int t0, int ts, F &&body)
{
  int k = k0 + t0;
  int cur = k < k1 ? src[k] : 0;
  aw_hold(cur);  // complete on entry as on the back edge: no wait in the loop
  for (; k < k1; k += ts) {
    int nxt = src[k + ts < k1 ? k + ts : k1 - 1];
    body(cur >> 16, cur & 0xffff, nxt);
    cur = nxt;
  }
}

template <class T, class F>
AW_DEV void for_pairs_range(const View<T> &V, int k0, int k1, int t0, int ts, F &&body)
{
  const AW_G int *lst = (V.nl && V.nl_state[7] > 0) ? V.nl + AW_NL_HDR : nullptr;
  if (lst) {
    for_packed(lst, k0, k1, t0, ts, body);
    return;
  }
  int n = V.n;
  for (int k = k0 + t0; k < k1; k += ts) {
    int i, j;
    pair_decode(k, n, i, j);
    body(i, j);
  }
}

template <class T, class F>
AW_DEV void for_pairs(const View<T> &V, int t0, int ts, F &&body)
{
  for_pairs_range(V, 0, pair_total(V), t0, ts, body);
}

// pairs for the water-xi and force passes: the density-pair queue filled by
// phase B's list scan (resident engine, valid list), else every pair
template <class T, class F>
AW_DEV void for_density_pairs(const View<T> &V, int t0, int ts, F &&body)
{
  // body(i, j, queued): queued pairs passed the site-distance test of the
  // scan, which implies the near_pair test (|CA-CA| <= |site-site| + extents)
  const AW_G int *qs = (V.wq && V.nl && V.nl_state[7] > 0) ? V.wq + 4 * V.nl_cap : nullptr;
  if (qs) {
    for_packed(qs, 0, V.nl_state[5], t0, ts, [&](int i, int j) { body(i, j, true); });
    return;
  }
  for_pairs(V, t0, ts, [&](int i, int j) { body(i, j, false); });
}

// ------------------------------------------------------------------------
// Phase B

template <class T>
AW_DEV void residue_local(const View<T> &V, int i, EAcc &E)
{
  ParamPtr P = V.P;
  int ri = V.rinfo[i];
  bool first = ri & RI_FIRST, last = ri & RI_LAST, gly = ri & RI_GLY;
  // the switches, read before any branch
  const unsigned mask = P->term_mask;
  const bool chain_on = (P->chain_flag != 0) & ((mask >> AW_ET_CHAIN) & 1u);
  const bool chi_on = (P->chi_flag != 0) & ((mask >> AW_ET_CHI) & 1u);
  const bool shake_on = (P->shake_flag != 0) & ((mask >> AW_ET_SHAKE) & 1u);
  const bool rama_on = (P->rama_flag != 0) & ((mask >> AW_ET_RAMA) & 1u);
  const bool bond_on = (P->bond_on != 0) & ((mask >> AW_ET_BOND) & 1u);
  const bool ev_on = (P->ev_on != 0) & ((mask >> AW_ET_VEXCLUDED) & 1u);
  LocF F;
  F.clear();
  if (chain_on) chain_term(V, i, ri, E, F);
  if (!first && !last && chi_on && !gly) chi_term(V, i, E, F);
  if (shake_on) shake_term(V, i, ri, E, F);
  if (!first && !last && rama_on && !gly) rama_term(V, i, ri, E, F);
  if (bond_on) bonds_term(V, i, ri, E, F);
  if (ev_on) ev_self(V, i, ri, E, F);
  F.flush(V, i, first, last);
}

// Residue-pair pass of phase B.  Q (queues qd/qp/qe given): the DSSP
// (directed, acceptor << 16 | donor), P_AP and excluded-volume pairs that
// pass their cutoffs are queued and evaluated afterwards in dense passes;
// otherwise they are evaluated in place.  A template argument, so the scan
// loop of the queued pass carries no in-place term code.
// where the list scan adds the densities rho_w (w = 0) and rho_h (w = 1):
// the slab's per-residue scalars, or (DensWin) a workgroup's LDS window of
// the two, rho_w of residue i at win[2 i], rho_h at win[2 i + 1]
struct DensAdd {
  template <class VT>
  AW_DEV void operator()(const VT &V, int w, int i, double v) const { aadd(w == 0 ? &V.rho_w()[i] : &V.rho_h()[i], v); }
};
struct DensWin {
  double *win;
  template <class VT>
  AW_DEV void operator()(const VT &, int w, int i, double v) const { aadd(win + 2 * i + w, v); }
};

template <class T, bool Q = false, class DS = DensAdd>
AW_DEV void pair_B(const View<T> &V, int i, int j, EAcc &E, AW_G int *qd = nullptr, AW_G int *qp = nullptr,
                   AW_G int *qe = nullptr, AW_G int *qs = nullptr, int *hold = nullptr, DS dens = DS())
{
  ParamPtr P = V.P;
  // every parameter the scan reads, loaded before any branch: invariant and
  // unconditional, so the loads leave the pair loop (conditional ones would
  // be reissued, and waited for, every iteration)
  const float reach = (float)P->pair_reach, dreach = (float)P->dens_reach, evreach = (float)P->ev_reach;
  const bool water = P->water_flag, helix = P->helix_flag;
  const double wmin2 = P->w_rmin_th_sq[0], wmax2 = P->w_rmax_th_sq[0], hmin2 = P->h_rmin_th_sq, hmax2 = P->h_rmax_th_sq;
  const double wk = P->water_kappa, wr0 = P->well_r_min[0], wr1 = P->well_r_max[0];
  const double hk = P->helix_kappa, hr0 = P->helix_r_min, hr1 = P->helix_r_max;
  const unsigned mask = P->term_mask;  // (no short-circuit: every load unconditional)
  const bool dssp_on = (P->dssp_flag != 0) & ((mask >> AW_ET_DSSP) & 1u);
  const bool pap_on = (P->pap_flag != 0) & ((mask >> AW_ET_PAP) & 1u);
  const bool ev_on = (P->ev_on != 0) & ((mask >> AW_ET_VEXCLUDED) & 1u);
  const double dcut = P->dssp_cut, dcut2 = P->dssp_cut_sq, pcut2 = P->pap_cutoff_sq;
  // CA-CA distance (FP32) against each term's reach (the near_pair test)
  // the pair's first reads (CA positions, extents, residue info) issued
  // together, ahead of any use: one LDS round trip instead of one per
  // quantity
  const double *pci = XA(V, K_CA, i), *pcj = XA(V, K_CA, j);
  const double ci[3] = {pci[0], pci[1], pci[2]}, cj[3] = {pcj[0], pcj[1], pcj[2]};
  const float exi = V.ext()[i], exj = V.ext()[j];
  const int ri0 = V.rinfo[i], rj0 = V.rinfo[j];
  AW_SCHED_FENCE();
  float rca = rsq_f(ci, cj);
  float eij = exi + exj + 1e-3f;
  float lim = reach + eij;
  // no early exit for pairs beyond every reach: all lanes reach the queue
  // stores through one join (a divergent return is joined after the
  // stores, and its path would then wait for them)
  const bool near = rca <= lim * lim * 1.0001f;
  bool dq = false;  // a pair the water-xi and force passes (phases D, F) can use
  bool dij = false, dji = false, pap = false, ev = false;
  int ri = 0, rj = 0;
  if (near) {
    ri = ri0; rj = rj0;
    bool same = rchain(ri) == rchain(rj);
    int dres = j - i;
    // DSSP in both orientations (fb.cpp:8064-8094); ahead of the density
    // block, whose atomic adds would otherwise order these reads behind them
    float ld = (float)dcut + eij;
    if (dssp_on && (!same || dres > 2) && rca <= ld * ld * 1.0001f) {
      if (!(rj & RI_PRO) && !(ri & RI_LAST) && !(rj & RI_FIRST))
        dij = dist_within(XA(V, K_O, i), XV(V, K_N, j), dcut, dcut2);
      if (!(ri & RI_PRO) && !(rj & RI_LAST) && !(ri & RI_FIRST))
        dji = dist_within(XA(V, K_O, j), XV(V, K_N, i), dcut, dcut2);
    }
    // DL1 densities (fb.cpp:7729-7786); the exact FP64 distance only for
    // pairs near a well
    if (!same || dres > 1) {
      const double *si = site(V, i, ri), *sj = site(V, j, rj);
      float e = rsq_f(si, sj);
      if (qs) dq = e < dreach * dreach * (1.0f + AW_RSQ_MARGIN);
      bool wc = water && e > (float)wmin2 * (1.0f - AW_RSQ_MARGIN) && e < (float)wmax2 * (1.0f + AW_RSQ_MARGIN);
      bool hc = helix && e > (float)hmin2 * (1.0f - AW_RSQ_MARGIN) && e < (float)hmax2 * (1.0f + AW_RSQ_MARGIN);
      if (wc || hc) {
        double rsq = rsq_exact(si, sj);
        if (wc && rsq > wmin2 && rsq < wmax2) {
          T th, prd;
          well_theta<T>(wk, wr0, wr1, sqrt_t<T>(rsq), th, prd);
          dens(V, 0, i, (double)th);
          dens(V, 0, j, (double)th);
        }
        if (hc && rsq > hmin2 && rsq < hmax2) {
          T th, prd;
          well_theta<T>(hk, hr0, hr1, sqrt_t<T>(rsq), th, prd);
          dens(V, 1, i, (double)th);
          dens(V, 1, j, (double)th);
        }
      }
    }
    // P_AP (fb.cpp:8096-8103)
    pap = pap_on && rsq_lt_est(rca, ci, cj, pcut2);
    float le = evreach + eij;
    ev = ev_on && rca <= le * le * 1.0001f;
  }
  if constexpr (Q) {
    if (hold) aw_hold(*hold);  // the list scan's next entry, before the stores
    // queue counters nl_state[2..5]: DSSP (both orientations), P_AP,
    // excluded volume, density pairs
    lmask m1 = w_ballot(dij), m2 = w_ballot(dji), mp = w_ballot(pap), me = w_ballot(ev), ms = w_ballot(dq);
    int n1 = w_count(m1);
    int cnt[4] = {n1 + w_count(m2), w_count(mp), w_count(me), w_count(ms)}, base[4];
    q_reserve4(&V.nl_state[2], cnt, base);
    int c = (i << 16) | j;
    if (dij) qd[base[0] + w_rank(m1)] = c;
    if (dji) qd[base[0] + n1 + w_rank(m2)] = (j << 16) | i;
    if (pap) qp[base[1] + w_rank(mp)] = c;
    if (ev) qe[base[2] + w_rank(me)] = c;
    if (dq) qs[base[3] + w_rank(ms)] = c;
  } else {
    if (dij) dssp_term(V, i, j, E);
    if (dji) dssp_term(V, j, i, E);
    if (pap) pap_term(V, i, j, E);
    if (ev) ev_pair(V, i, j, ri, rj, E);
  }
}

// fragment-memory table mode (table_fragment_memory, fb.cpp:4325-4429 over
// the pair path's loop 7513-7527), one item = one atom pair (i, j, q, c) of
// that loop (q: 0 CA-CA, 1 CA(i)-CB(j), 2 CB(i)-CA(j), 3 CB-CB; glycine
// skips and the atom pairs after a missing curve are left out when the
// items are built): the FP64 distance r must lie in [rmin, rmax], else the
// error word is set (the reference aborts); c < 0 (a curve of zeros or a
// missing one) is a range check only; otherwise energy and force factor are
// interpolated linearly between the grid points around r, and
// f_i += ff dx, f_j -= ff dx.  Items k, k + ks, ... (< k1) of one lane,
// TB_NB at a time: their curve loads are issued together, so a lane waits
// for one memory latency per batch instead of one per atom pair.
// Coordinates x[3][n][3], forces at fb + 3 (kind fs + i) (the view's force
// base and stride); returns the energy.  A function of its own (not
// inlined): table mode is rare, and its registers stay out of the passes
// around its call.
struct alignas(16) TbItem { int i, j, q, c; };

template <class T>
AW_NOINLINE double fm_table_run(const double *x, int n, double *fb, int fs, ParamPtr P, const AW_G int *items,
                                const void *curve, AW_G int *fault, int k, int k1, int ks)
{
  typedef typename std::conditional<sizeof(T) == 4, float2, double2>::type T2;
  constexpr int TB_NB = 4;
  const double rmin = P->tb_rmin, rmax = P->tb_rmax, dr = P->tb_dr;
  const int tbs = P->tb_size;
  const AW_G TbItem *it4 = (const AW_G TbItem *)items;
  const T2 *cv2 = (const T2 *)curve;
  double e = 0.0;
  int bad = 0;
  for (; k < k1; k += TB_NB * ks) {
    int ia[TB_NB], ja[TB_NB];
    double d[TB_NB][3];
    T tw[TB_NB];
    T2 lo[TB_NB], hi[TB_NB];
    bool use[TB_NB];
#pragma unroll
    for (int u = 0; u < TB_NB; ++u) {
      int kk = k + u * ks;
      use[u] = false;
      if (kk < k1) {
        const AW_G TbItem &w = it4[kk];
        int wi = w.i, wj = w.j, wq = w.q, wc = w.c;
        int ka = wq >> 1, kb = wq & 1;  // K_CA = 0, K_CB = 1
        const double *a = x + (ka * n + wi) * 3, *b = x + (kb * n + wj) * 3;
        d[u][0] = a[0] - b[0]; d[u][1] = a[1] - b[1]; d[u][2] = a[2] - b[2];
        double r = sqrt(d[u][0] * d[u][0] + d[u][1] * d[u][1] + d[u][2] * d[u][2]);
        ia[u] = ka * fs + wi;
        ja[u] = kb * fs + wj;
        if (!(r >= rmin && r <= rmax)) {
          bad = 1;
        } else if (wc >= 0) {
          int ir = (int)((r - rmin) / dr);
          ir = ir < tbs - 2 ? ir : tbs - 2;
          tw[u] = (T)((r - (rmin + (double)ir * dr)) / dr);
          const T2 *cv = cv2 + (size_t)wc * tbs + ir;
          lo[u] = cv[0];
          hi[u] = cv[1];
          use[u] = true;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < TB_NB; ++u) {
      if (!use[u]) continue;
      T ve = lo[u].x + (hi[u].x - lo[u].x) * tw[u], ff = lo[u].y + (hi[u].y - lo[u].y) * tw[u];
      e += (double)ve;
      double fx = (double)(ff * (T)d[u][0]), fy = (double)(ff * (T)d[u][1]), fz = (double)(ff * (T)d[u][2]);
      double *pi = fb + ia[u] * 3, *pj = fb + ja[u] * 3;
      aadd(pi + 0, fx); aadd(pi + 1, fy); aadd(pi + 2, fz);
      aadd(pj + 0, -fx); aadd(pj + 1, -fy); aadd(pj + 2, -fz);
    }
  }
#ifndef AW_HOST_EMU
  if (bad) atomicOr((int *)fault, AW_FAULT_TB_RANGE);
#endif
  return e;
}

// the table pass over one replica's items by the lanes t0, t0 + ts, ...;
// its energy into the fragment-memory accumulator when reported (ef)
template <class T>
AW_DEV void fm_table_pass(const View<T> &V, int t0, int ts, bool ef)
{
  ParamPtr P = V.P;
  if (!(P->fm_tb_flag && ON(P, AW_ET_FRAGMEM) && V.tb_items > 0)) return;
  double e = fm_table_run<T>(V.x(), V.n, V.fb, V.fs, P, V.tb_item, V.tb_curve, V.fault, t0, V.tb_items, ts);
  if (ef) {
    EAcc E;
    E.clear();
    E.fm = e;
    eflush(E, V.ered);
  }
}

// part: bit 0 fragment memory, bit 1 local terms, bit 2 residue pairs.
// QUEUED (resident engine, one workgroup per replica): the pair sweep runs
// in chunks of at most nl_cap pairs; each chunk is scanned into the DSSP,
// P_AP and excluded-volume work queues, which are then evaluated densely.
// Unqueued (grid-wide phase engine): every term is evaluated in place.
// pm (clocked runs only): pm->mark(k) after each barrier of the queued pair
// pass, k = 12 scan, 13 DSSP, 14 P_AP, 15 excluded volume
// XLDS: the working slab (coordinates, forces) is in LDS; ef: this
// evaluation's energies are reported (fragment memory skips its energy sums
// otherwise); TB: the fragment-memory table pass is compiled in (an
// instance without it has no call in its pair pass)
template <class T, bool QUEUED, class PM = void, bool XLDS = false, bool TB = true>
AW_DEV void phase_B(const View<T> &V, int t0, int ts, int part, PM *pm = nullptr, bool ef = true)
{
  EAcc E;
  E.clear();
  ParamPtr P = V.P;
  int nfm = (P->fm_flag && ON(P, AW_ET_FRAGMEM) && (part & 1)) ? V.fm_items : 0;
  auto run_fm = [&]() {
    if (nfm <= 0) return;
    typedef typename T4<T>::item ItemT;
    const ItemT *items = (const ItemT *)V.items + V.fm_item_base;
    // the queued (resident) pass gives a lane many items: pipelined loads
    typedef FmSinkForce<XLDS> Sk;
    if (QUEUED && V.sh_n > 0) {
      E.fm += fm_claims<T, XLDS>(V, items, ef);
      return;
    }
    if (ef)
      E.fm += fm_run<T, XLDS, true, Sk, QUEUED>(V.x(), V.n, V.rf, V.coef, items, t0, nfm, ts, Sk{aw_cast<XLDS>(V.f()), V.n});
    else
      fm_run<T, XLDS, false, Sk, QUEUED>(V.x(), V.n, V.rf, V.coef, items, t0, nfm, ts, Sk{aw_cast<XLDS>(V.f()), V.n});
  };
  // fragment memory runs after the last chunk's queue passes, before their
  // closing barrier: a wave that is done with its queue items starts on
  // the arithmetic-bound fragment memory while the other waves of its SIMD
  // still wait on the queue passes' memory accesses (no pass reads what
  // another one writes: they only add forces)
  bool close = false;  // the last chunk's queue passes still to be closed
  if (part & 2)
    for (int k = t0; k < V.n; k += ts) residue_local(V, k, E);
  if (part & 4) {
    if (QUEUED) {
      int cap = V.nl_cap;
      AW_G int *qd = V.wq, *qp = V.wq + 2 * cap, *qe = V.wq + 3 * cap;
      int total = pair_total(V);
      for (int c0 = 0; c0 < total && cap > 0; c0 += cap) {
        int c1 = total - c0 < cap ? total : c0 + cap;
        // density pairs are kept for phases D and F when the list holds
        // every pair (one chunk)
        AW_G int *qs = total <= cap && V.nl && V.nl_state[7] > 0 ? V.wq + 4 * cap : nullptr;
        if (V.nl && V.nl_state[7] > 0)
          for_packed_st(V.nl + AW_NL_HDR, c0, c1, t0, ts,
                        [&](int i, int j, int &nx) { pair_B<T, true>(V, i, j, E, qd, qp, qe, qs, &nx); });
        else
          for_pairs_range(V, c0, c1, t0, ts, [&](int i, int j) { pair_B<T, true>(V, i, j, E, qd, qp, qe, qs); });
        AW_BARRIER();
        if constexpr (!std::is_void<PM>::value) if (pm) pm->mark(12);
        int nd = V.nl_state[2], npap = V.nl_state[3], nev = V.nl_state[4];
        // the queue lengths of a one-chunk pass into the list header (slots
        // 12-15), where awsem_batch_lists reads them after the launch
        if (t0 == 0 && qs) {
          V.nl[12] = nd; V.nl[13] = npap; V.nl[14] = nev; V.nl[15] = V.nl_state[5];
        }
        for_packed(qd, 0, nd, t0, ts, [&](int i, int j) { dssp_term(V, i, j, E); });
        if constexpr (!std::is_void<PM>::value) if (pm) { AW_BARRIER(); pm->mark(13); }
        for_packed(qp, 0, npap, t0, ts, [&](int i, int j) { pap_term(V, i, j, E); });
        if constexpr (!std::is_void<PM>::value) if (pm) { AW_BARRIER(); pm->mark(14); }
        for_packed(qe, 0, nev, t0, ts, [&](int i, int j) { ev_pair(V, i, j, V.rinfo[i], V.rinfo[j], E); });
        if (c0 + cap >= total) {
          close = true;  // closed after the fragment memory below
          break;
        }
        AW_BARRIER();
        if constexpr (!std::is_void<PM>::value) if (pm) pm->mark(15);
        if (t0 == 0) V.nl_state[2] = V.nl_state[3] = V.nl_state[4] = 0;
        AW_BARRIER();
      }
    } else {
      for_pairs(V, t0, ts, [&](int i, int j) { pair_B(V, i, j, E); });
    }
  }
  // the energies so far are flushed before the fragment-memory call, so
  // that no accumulator stays live across it; only an evaluation whose
  // energies are reported (ef) reduces them at all
  if (ef) eflush(E, V.ered);
  E.clear();
  run_fm();
  if (TB && P->fm_tb_flag && ON(P, AW_ET_FRAGMEM) && (part & 1) && V.tb_items > 0)
    E.fm += fm_table_run<T>(V.x(), V.n, V.fb, V.fs, P, V.tb_item, V.tb_curve, V.fault, t0, V.tb_items, ts);
  if (QUEUED && close) {
    AW_BARRIER();
    if constexpr (!std::is_void<PM>::value) if (pm) pm->mark(15);
    if (t0 == 0) V.nl_state[2] = V.nl_state[3] = V.nl_state[4] = 0;
    AW_BARRIER();
  }
  if (ef) eflush(E, V.ered);
}

// ------------------------------------------------------------------------
// Phase C: sigma_h and burial (SL, fb.cpp:7795-7841)

// sigma = (1 - tanh x) / 2 and 1 + tanh x (fb.cpp:7802-7804): FP64, or FP32
// through E = exp(-2|x|) without cancellation: sigma = E/(1+E) (x >= 0) or
// 1/(1+E), 1 + tanh = 2/(1+E) or 2E/(1+E)
template <class T>
AW_DEV void sig_switch(double x, double &s, double &onep)
{
  if (sizeof(T) == 8) {
    double th = tanh(x);
    s = 0.5 * (1.0 - th);
    onep = 1.0 + th;
    return;
  }
  float e = aw_fast_exp(-2.0f * (float)(x < 0.0 ? -x : x));
  float ip = aw_frcp(1.0f + e);
  s = (double)(x >= 0.0 ? e * ip : ip);
  onep = (double)(x >= 0.0 ? 2.0f * ip : 2.0f * e * ip);
}
// tanh(x) and sech^2(x) = 1 - tanh^2(x) in FP32 through E = exp(-2|x|):
// tanh = +-(1-E)/(1+E), sech^2 = 4E/(1+E)^2 (no cancellation for |x| large)
template <class T>
AW_DEV void sw_tanh_sech2(double x, double &t, double &s)
{
  float e = aw_fast_exp(-2.0f * (float)(x < 0.0 ? -x : x));
  float ip = aw_frcp(1.0f + e);
  float tt = (1.0f - e) * ip;
  t = x < 0.0 ? -(double)tt : (double)tt;
  s = (double)(4.0f * e * ip * ip);
}

template <class T>
AW_DEV void phase_C(const View<T> &V, int t0, int ts, bool ef = true)
{
  EAcc E;
  E.clear();
  ParamPtr P = V.P;
  for (int i = t0; i < V.n; i += ts) {
    // a phosphorylated residue takes type 6's burial gammas (fb.cpp:2887-2900)
    int it = (V.rinfo[i] & RI_PHOS) ? 6 : rtype(V.rinfo[i]);
    // the residue's reads before its stores (which could alias them) and
    // before any branch: densities and its burial gammas (per-lane gathers)
    const double rho_w = V.rho_w()[i], rho_h = V.rho_h()[i];
    const double bg[3] = {P->burial_gamma[it][0], P->burial_gamma[it][1], P->burial_gamma[it][2]};
    if (P->water_flag) {
      double s, onep;
      sig_switch<T>(P->water_kappa_sigma * (rho_w - P->treshold), s, onep);
      V.sig_w()[i] = s;
      V.sigp_w()[i] = -P->water_kappa_sigma * s * onep;
    }
    if (P->helix_flag) {
      double s, onep;
      sig_switch<T>(P->helix_kappa_sigma * (rho_h - P->helix_treshold), s, onep);
      V.sig_h()[i] = s;
      V.sigp_h()[i] = -P->helix_kappa_sigma * s * onep;
    }
    if (P->burial_flag) {
      double rho = rho_w, kb = P->burial_kappa, bfv = 0.0, eb = 0.0;
      for (int b = 0; b < 3; ++b) {
        double x0 = kb * (rho - P->burial_ro_min[b]), x1 = kb * (P->burial_ro_max[b] - rho);
        double t0_, t1_, dsq;
        if (sizeof(T) == 8) {
          t0_ = tanh(x0);
          t1_ = tanh(x1);
          dsq = t1_ * t1_ - t0_ * t0_;
        } else {
          // t1^2 - t0^2 = sech^2(x0) - sech^2(x1), with sech^2 = 4E/(1+E)^2,
          // E = exp(-2|x|): exact to FP32 rounding relative to the sech^2
          // values, where 1 - t^2 formed from t near +-1 would cancel
          double s0, s1;
          sw_tanh_sech2<T>(x0, t0_, s0);
          sw_tanh_sech2<T>(x1, t1_, s1);
          dsq = s0 - s1;
        }
        eb += -0.5 * P->k_burial * bg[b] * (t0_ + t1_);
        bfv += bg[b] * dsq;
      }
      V.bf()[i] = bfv * 0.5 * P->k_burial * kb;
      if (ON(P, AW_ET_BURIAL)) E.burial += eb;
    }
  }
  if (ef) eflush(E, V.ered);  // energies only where reported
}

// ------------------------------------------------------------------------
// Phase D: helix pairs (fb.cpp:7843-7868, 3706-3785) and water xi (DL2)

template <class T>
AW_DEV void helix_item(const View<T> &V, int i, EAcc &E)
{
  ParamPtr P = V.P;
  int j = i + P->helix_i_diff;
  if (j >= V.n) return;
  int ri = V.rinfo[i], rj = V.rinfo[j];
  if (rchain(ri) != rchain(rj)) return;
  double r2 = rsq_exact(XA(V, K_O, i), XV(V, K_N, j));
  if (!(r2 < P->helix_cutoff_sq)) return;
  if (sqrt(r2) > P->helix_cutoff) return;
  V3<T> xNO = dvec<T>(XA(V, K_O, i), XV(V, K_N, j)), xHO = dvec<T>(XA(V, K_O, i), XV(V, K_H, j));
  double RNOd = sqrt(r2), RHOd = sqrt(rsq_exact(XA(V, K_O, i), XV(V, K_H, j)));
  T RNO = (T)RNOd, RHO = (T)RHOd;
  int it = rtype(ri), jt = rtype(rj);
  double h4i = ((ri & RI_PRO) && P->pro_accepter_flag) ? P->h4prob_pro_accepter : P->h4prob[it];
  T prob_sum = (T)(h4i + P->h4prob[jt]);
  T dNO = (T)(RNOd - P->helix_NO_zero), dHO = (T)(RHOd - P->helix_HO_zero);
  T dNOs = dNO * (T)P->helix_sigma_NO_sqinv, dHOs = dHO * (T)P->helix_sigma_HO_sqinv;
  T pt = -(T)P->k_helix * prob_sum * t_exp((T)-0.5 * (dNO * dNOs + dHO * dHOs));
  T p0 = t_div(-dNOs, RNO), p1 = t_div(-dHOs, RHO);
  double gdiff = P->helix_gamma_w - P->helix_gamma_p;
  double sg = P->helix_gamma_p + gdiff * V.sig_h()[i] * V.sig_h()[j];
  V.hxi1()[i] = gdiff * (double)pt * V.sigp_h()[i] * V.sig_h()[j];
  V.hxi2()[i] = gdiff * (double)pt * V.sig_h()[i] * V.sigp_h()[j];
  if (!ON(P, AW_ET_HELIX)) return;
  T Vv = (T)sg * pt;
  E.helix += (double)Vv;
  T an = (T)P->an, bn = (T)P->bn, cn = (T)P->cn, ah = (T)P->ah, bh = (T)P->bh, ch = (T)P->ch;
  T q0 = Vv * p0, q1 = Vv * p1;
  fadd3(V, K_CA, j - 1, an * q0 * xNO.x + ah * q1 * xHO.x, an * q0 * xNO.y + ah * q1 * xHO.y, an * q0 * xNO.z + ah * q1 * xHO.z);
  fadd3(V, K_CA, j, bn * q0 * xNO.x + bh * q1 * xHO.x, bn * q0 * xNO.y + bh * q1 * xHO.y, bn * q0 * xNO.z + bh * q1 * xHO.z);
  fadd3(V, K_O, j - 1, cn * q0 * xNO.x + ch * q1 * xHO.x, cn * q0 * xNO.y + ch * q1 * xHO.y, cn * q0 * xNO.z + ch * q1 * xHO.z);
  fadd3(V, K_O, i, -(q0 * xNO.x + q1 * xHO.x), -(q0 * xNO.y + q1 * xHO.y), -(q0 * xNO.z + q1 * xHO.z));
}

// one water well's thresholds and range (wells 0 and 1, the shipped deck's,
// are read before the pair loops; further wells inside them)
struct WellP {
  double min2, max2, r0, r1;
  int flag;
};
AW_DEV WellP well_p(ParamPtr P, int w)
{
  WellP q;
  q.min2 = P->w_rmin_th_sq[w]; q.max2 = P->w_rmax_th_sq[w];
  q.r0 = P->well_r_min[w]; q.r1 = P->well_r_max[w];
  q.flag = P->well_flag[w];
  return q;
}

// where the water-xi pass adds xi_w: the slab's per-residue scalar, or
// (XiWin) a workgroup's LDS window, xi_w of residue i at win[i]
struct XiAdd {
  template <class VT>
  AW_DEV void operator()(const VT &V, int i, double v) const { aadd(&V.xi_w()[i], v); }
};
struct XiWin {
  double *win;
  template <class VT>
  AW_DEV void operator()(const VT &, int i, double v) const { aadd(win + i, v); }
};

template <class T, class XS = XiAdd>
AW_DEV void xi_pair(const View<T> &V, int i, int j, bool queued, XS xs = XS())
{
  ParamPtr P = V.P;
  // parameters read before any branch (invariant loads leave the loop)
  const int cc = P->contact_cutoff, nw = P->n_wells;
  const float dreach = (float)P->dens_reach;
  const double wk = P->water_kappa;
  const WellP w0 = well_p(P, 0), w1 = well_p(P, 1);
  int ri = V.rinfo[i], rj = V.rinfo[j];
  if (!(rchain(ri) != rchain(rj) || j - i >= cc)) return;
  if (!queued && !near_pair(V, i, j, dreach)) return;
  int it = rtype(ri), jt = rtype(rj);
  // wells 0 and 1's gammas (per-lane gathers), issued before the distance
  // so that their latency overlaps it
  // a pair with a phosphorylated residue reads the hypercharged table (fb.cpp:2872-2885)
  const auto wg = ((ri | rj) & RI_PHOS) ? P->phosph_gamma : P->water_gamma;
  const double g00 = wg[0][it][jt][0], g01 = wg[0][it][jt][1];
  const double g10 = wg[1][it][jt][0], g11 = wg[1][it][jt][1];
  const double swi = V.sig_w()[i], swj = V.sig_w()[j];  // with the positions
  double rsq = rsq_exact(site(V, i, ri), site(V, j, rj));
  double r = sqrt_t<T>(rsq);
  double acc_i = 0.0, acc_j = 0.0;
  for (int w = 0; w < nw; ++w) {
    WellP wp = w == 0 ? w0 : (w == 1 ? w1 : well_p(P, w));
    if (!wp.flag) continue;
    double g0 = w == 0 ? g00 : (w == 1 ? g10 : wg[w][it][jt][0]);
    double g1 = w == 0 ? g01 : (w == 1 ? g11 : wg[w][it][jt][1]);
    // fb.cpp:7925 reads !fabs(g0-g1)<delta, i.e. (!fabs(g0-g1)) < delta: true iff g0 != g1
    if (((double)(!fabs(g0 - g1)) < AW_DELTA) && rsq > wp.min2 && rsq < wp.max2) {
      T th, prd;
      well_theta<T>(wk, wp.r0, wp.r1, r, th, prd);
      double tg = (g1 - g0) * (double)th;
      acc_i += tg * swj;
      acc_j += tg * swi;
    }
  }
  if (acc_i != 0.0) xs(V, i, acc_i);
  if (acc_j != 0.0) xs(V, j, acc_j);
}

template <class T, class XS = XiAdd>
AW_DEV void phase_D(const View<T> &V, int t0, int ts, bool ef = true, XS xs = XS())
{
  EAcc E;
  E.clear();
  ParamPtr P = V.P;
  int n = V.n;
  int nh = P->helix_flag ? n : 0;
  for (int k = t0; k < nh; k += ts) helix_item(V, k, E);
  if (P->water_flag) for_density_pairs(V, t0, ts, [&](int i, int j, bool q) { xi_pair(V, i, j, q, xs); });
  if (ef) eflush(E, V.ered);  // energies only where reported
}

// ------------------------------------------------------------------------
// Phase F: DL3 density-site pair forces (fb.cpp:7996-8057, scatter 8116-8128)

template <class T>
AW_DEV void force_pair(const View<T> &V, int i, int j, EAcc &E, bool queued)
{
  ParamPtr P = V.P;
  // parameters read before any branch (invariant loads leave the loop)
  const unsigned mask = P->term_mask;
  const bool water = (P->water_flag != 0) & ((mask >> AW_ET_WATER) & 1u);
  const bool burial = (P->burial_flag != 0) & ((mask >> AW_ET_BURIAL) & 1u);
  const bool helix = (P->helix_flag != 0) & ((mask >> AW_ET_HELIX) & 1u);
  const int cc = P->contact_cutoff, nw = P->n_wells, hd = P->helix_i_diff;
  const float dreach = (float)P->dens_reach;
  const double wk = P->water_kappa, hk = P->helix_kappa, hr0 = P->helix_r_min, hr1 = P->helix_r_max;
  const double hmin2 = P->h_rmin_th_sq, hmax2 = P->h_rmax_th_sq;
  const WellP w0 = well_p(P, 0), w1 = well_p(P, 1);
  int ri = V.rinfo[i], rj = V.rinfo[j];
  bool same = rchain(ri) == rchain(rj);
  int dres = j - i;
  if (!(!same || dres > 1)) return;
  if (!queued && !near_pair(V, i, j, dreach)) return;
  int it = rtype(ri), jt = rtype(rj);
  // wells 0 and 1's gammas (per-lane gathers), issued before the distance
  // so that their latency overlaps it
  const auto wg = ((ri | rj) & RI_PHOS) ? P->phosph_gamma : P->water_gamma;
  const double g00 = wg[0][it][jt][0], g01 = wg[0][it][jt][1];
  const double g10 = wg[1][it][jt][0], g11 = wg[1][it][jt][1];
  const double *si = site(V, i, ri), *sj = site(V, j, rj);
  // the per-residue scalars the pair can use, read with the positions and
  // before any branch (one LDS round trip; hxi2 indices clamped, their
  // values used only where i - hd, j - hd >= 0)
  const double swi = V.sig_w()[i], swj = V.sig_w()[j];
  const double xii = V.xi_w()[i] * V.sigp_w()[i], xij = V.xi_w()[j] * V.sigp_w()[j];
  const double bfij = V.bf()[i] + V.bf()[j];
  const double h1 = V.hxi1()[i] + V.hxi1()[j];
  const double h2i = V.hxi2()[i - hd >= 0 ? i - hd : 0], h2j = V.hxi2()[j - hd >= 0 ? j - hd : 0];
  double rsq = rsq_exact(si, sj);
  double r = sqrt_t<T>(rsq);
  T force = 0;
  bool in0 = rsq > w0.min2 && rsq < w0.max2;
  // well 0's theta'/r, shared by the water, water-xi and burial parts
  T th0 = 0, prd0 = 0;
  if (in0) well_theta<T>(wk, w0.r0, w0.r1, r, th0, prd0);
  if (water) {
    double ew = 0.0;
    for (int w = 0; w < nw; ++w) {
      WellP wp = w == 0 ? w0 : (w == 1 ? w1 : well_p(P, w));
      if (!wp.flag) continue;
      if ((!same || dres >= cc) && rsq > wp.min2 && rsq < wp.max2) {
        double g0 = w == 0 ? g00 : (w == 1 ? g10 : wg[w][it][jt][0]);
        double g1 = w == 0 ? g01 : (w == 1 ? g11 : wg[w][it][jt][1]), sg;
        if (fabs(g0 - g1) < AW_DELTA) sg = 0.5 * (g0 + g1);
        else sg = g0 + (g1 - g0) * swi * swj;
        T th = th0, prd = prd0;
        if (w != 0) well_theta<T>(wk, wp.r0, wp.r1, r, th, prd);
        ew += -sg * (double)th;
        force += (T)sg * prd;
      }
    }
    E.water += ew;
    if (in0) {
      double xi = 0.0;
      // xii, xij: xi times dsigma/drho (fb.cpp:7939-7950), formed where it
      // is read: the same product, without a pass and a barrier of its own
      if (fabs(xii) > AW_DELTA_WATER_XI) xi += xii;
      if (fabs(xij) > AW_DELTA_WATER_XI) xi += xij;
      if (xi != 0.0) force += prd0 * (T)xi;
    }
  }
  if (burial && in0) force += (T)bfij * prd0;
  if (helix && rsq > hmin2 && rsq < hmax2) {
    double factor = h1;
    if (i - hd >= 0) factor += h2i;
    if (j - hd >= 0) factor += h2j;
    if (fabs(factor) > AW_DELTA_HELIX_XI) {
      T th, prd;
      well_theta<T>(hk, hr0, hr1, r, th, prd);
      force += -(T)factor * prd;
    }
  }
  if (force != (T)0) {
    V3<T> d = dvec<T>(si, sj);
    int ki = (ri & RI_GLY) ? K_CA : K_CB, kj = (rj & RI_GLY) ? K_CA : K_CB;
    fadd(V, ki, i, force, d);
    fadd(V, kj, j, -force, d);
  }
}

template <class T>
AW_DEV void phase_F(const View<T> &V, int t0, int ts, bool ef = true)
{
  EAcc E;
  E.clear();
  for_density_pairs(V, t0, ts, [&](int i, int j, bool q) { force_pair(V, i, j, E, q); });
  if (ef) eflush(E, V.ered);  // energies only where reported
}

// ------------------------------------------------------------------------
// Langevin dynamics (LAMMPS fix nve + fix langevin, real units)

#define AW_FTM2V (1.0 / 48.88821291 / 48.88821291)
#define AW_MVV2E (48.88821291 * 48.88821291)
#define AW_BOLTZ 0.0019872067

AW_DEV uint64_t mix64(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
// counter hash (seed, replica, step, atom, component) -> U[0,1), in two
// parts: the state after (seed, replica, component, step), then the atom
AW_DEV uint64_t urand_step(uint64_t seed, int replica, long long step, int comp)
{
  uint64_t h = mix64(seed ^ 0x9e3779b97f4a7c15ULL);
  h = mix64(h ^ ((uint64_t)(uint32_t)replica << 32 | (uint32_t)comp));
  return mix64(h ^ (uint64_t)step);
}
AW_DEV double urand_atom(uint64_t hstep, int atom)
{
  uint64_t h = mix64(hstep ^ (uint64_t)(uint32_t)atom);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
AW_DEV double urand(uint64_t seed, int replica, long long step, int atom, int comp)
{
  return urand_atom(urand_step(seed, replica, step, comp), atom);
}

// mass class of an atom: 0 CA (27), 1 CB (60), 2 O (28), 3 glycine H-beta (2)
AW_DEV int mass_class(int kind, int ri) { return (kind == K_CB && (ri & RI_GLY)) ? 3 : kind; }

// integrator constants; the per-class factors are the expressions the
// integrator would form per atom (same operations, same order), formed once
struct LgParams {
  double dt, damp;  // fs; fix langevin damping time (fs)
  double dtfm[4];   // 0.5 dt ftm2v / m
  double g1[4];     // -m / damp / ftm2v
  double g2c[4];    // sqrt(m) sqrt(24 kB / damp / dt / mvv2e) / ftm2v (times sqrt(T))
};
inline AW_HOSTDEV LgParams lg_make(double dt, double damp)
{
  LgParams lg;
  lg.dt = dt;
  lg.damp = damp;
  const double mass[4] = {27.0, 60.0, 28.0, 2.0};
  double dtf = 0.5 * dt * AW_FTM2V;
  for (int c = 0; c < 4; ++c) {
    double m = mass[c];
    lg.dtfm[c] = dtf / m;
    lg.g1[c] = -m / damp / AW_FTM2V;
    lg.g2c[c] = sqrt(m) * sqrt(24.0 * AW_BOLTZ / damp / dt / AW_MVV2E) / AW_FTM2V;
  }
  return lg;
}

// initial_integrate of fix nve
template <class T>
AW_DEV void integrate1(const View<T> &V, AW_G double *v, const LgParams &lg, int t0, int ts)
{
  int n = V.n;
  for (int a = t0; a < 3 * n; a += ts) {
    int kind = (a >= n) + (a >= 2 * n), i = a - kind * n;
    double dtfm = lg.dtfm[mass_class(kind, V.rinfo[i])];
    for (int l = 0; l < 3; ++l) {
      double vv = v[3 * a + l] + dtfm * V.f()[3 * a + l];
      v[3 * a + l] = vv;
      V.x()[3 * a + l] += lg.dt * vv;
    }
  }
}

// fix langevin post_force + fix nve final_integrate
template <class T>
AW_DEV void integrate2(const View<T> &V, AW_G double *v, const LgParams &lg, const ReplicaDesc &R, int s, int t0, int ts)
{
  int n = V.n;
  // the replica's run constants, read together ahead of any use (one
  // memory round trip)
  const double t_start = R.t_start, t_stop = R.t_stop;
  const unsigned long long seed = R.seed;
  const long long step0 = R.step0;
  const int ramp0 = R.ramp0, run_steps = R.run_steps, rng_id = R.rng_id;
  AW_SCHED_FENCE();
  // the target temperature ramps over the replica's whole run (which may
  // span several launches)
  double delta = run_steps > 0 ? (double)(ramp0 + s + 1) / (double)run_steps : 0.0;
  double Tt = t_start + delta * (t_stop - t_start);
  double tsq = sqrt(Tt);
  long long step = step0 + s;
  // urand's hash up to the step, per component (the atom enters last)
  uint64_t hs[3];
  for (int l = 0; l < 3; ++l) hs[l] = urand_step(seed, rng_id, step, l);
  for (int a = t0; a < 3 * n; a += ts) {
    int kind = (a >= n) + (a >= 2 * n), i = a - kind * n;
    int c = mass_class(kind, V.rinfo[i]);
    double g1 = lg.g1[c], g2 = lg.g2c[c] * tsq, dtfm = lg.dtfm[c];
    for (int l = 0; l < 3; ++l) {
      double fv = V.f()[3 * a + l] + g1 * v[3 * a + l] + g2 * (urand_atom(hs[l], a) - 0.5);
      V.f()[3 * a + l] = fv;
      v[3 * a + l] += dtfm * fv;
    }
  }
}

