// awsem_wide.h -- the wide engine: every phase is one launch over all
// replicas at once (blockIdx.y = replica, blockIdx.x spreads a replica's
// items over several workgroups), with each replica's working slab, Verlet
// list, work queues and their counters in global memory.  It keeps the whole
// GPU busy when replicas are few and large (multimer complexes), where one
// workgroup per replica (the resident engine) would leave most CUs idle.
//
// Step s of a launch: int1, A, list check/build (all pairs, or through a
// cell list for large replicas), list publish, B (fragment
// memory and local terms), B (list scan into the work queues), B (queues),
// C, D, E, F, int2.  Kernel boundaries are the barriers between phases.
#pragma once
#include <cstdlib>
#include "awsem_dev.h"

// replica of this block; false when it has no work in evaluation s
// (s = -1: the evaluation before the first step, skipped for replicas
// whose stored forces are valid; s = -2: a force evaluation on request)
__device__ __forceinline__ bool w_active(const ReplicaDesc &R, int s)
{
  if (s == -2) return true;
  if (s == -1) return !(R.flags & AW_RF_F_VALID) || R.nsteps == 0;
  return s < R.nsteps;
}

template <class T>
__device__ __forceinline__ View<T> w_view(const BatchDev &B, const ReplicaDesc &R)
{
  View<T> V = make_view<T>(B, R, R.scratch);
  V.nl = R.nl_cap > 0 ? (AW_G int *)R.nl : nullptr;
  V.nlx = (AW_G double *)R.nlx;
  V.nl_cap = R.nl_cap;
  V.nl_state = R.nl + 4;
  V.wq = V.nl ? V.nl + AW_NL_HDR + R.nl_cap : nullptr;
  return V;
}

#define W_TID (blockIdx.x * blockDim.x + threadIdx.x)
#define W_NT (gridDim.x * blockDim.x)

// launch start: state into the slab, list reset on new coordinates
template <class T>
__global__ void __launch_bounds__(256) k_w_begin(BatchDev B, int mode)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  View<T> V = w_view<T>(B, R);
  int n = R.n;
  bool have_f = mode == 1 && (R.flags & AW_RF_F_VALID) && R.nsteps > 0;
  for (int k = W_TID; k < 9 * n; k += W_NT) {
    V.x()[k] = R.x[k];
    if (have_f) V.f()[k] = R.f[k];
  }
  if (W_TID == 0 && V.nl) {
    if (R.flags & AW_RF_NL_RESET) V.nl[1] = 0;
    for (int k = 0; k < 8; ++k) V.nl_state[k] = 0;
    nl_state_init(V);
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_w_end(BatchDev B, int mode)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  View<T> V = w_view<T>(B, R);
  int n = R.n;
  for (int k = W_TID; k < 9 * n; k += W_NT) {
    if (mode == 1) R.x[k] = V.x()[k];
    R.f[k] = V.f()[k];
  }
  if (W_TID == 0) {
    double s = 0.0;
    for (int k = 1; k < AW_N_ET; ++k) s += R.energy[k];
    R.energy[0] = s;
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_w_A(BatchDev B, int s)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  if (!w_active(R, s)) return;
  View<T> V = w_view<T>(B, R);
  phase_A(V, W_TID, W_NT);
  if (W_TID < AW_N_ET) R.energy[W_TID] = 0.0;
}

// all-pairs list build (replicas without a cell table)
template <class T>
__global__ void __launch_bounds__(256) k_w_nl_build(BatchDev B, int s)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  if (!w_active(R, s) || R.cells) return;
  View<T> V = w_view<T>(B, R);
  if (nl_needs_build(V)) phase_nl_build(V, W_TID, W_NT);
}

// ---- cell-list build: residues binned by CA into cubic cells no smaller
// than the list reach (pair reach + skin + twice the largest extent), cells
// hashed into cell_h buckets; each residue then visits the 27 cells around
// its own and keeps the pairs near_pair accepts.  A pair is seen once: j > i
// and j's own cell must be the visited cell (buckets may mix cells).

__device__ __forceinline__ int w_cell_bucket(int cx, int cy, int cz, int h)
{
  unsigned u = (unsigned)cx * 73856093u ^ (unsigned)cy * 19349663u ^ (unsigned)cz * 83492791u;
  return (int)(u & (unsigned)(h - 1));
}

template <class T>
__device__ __forceinline__ float w_cell_size(const View<T> &V)
{
  float reach = (float)V.P->pair_reach + (float)V.P->nl_skin;
  return (reach + 2.0f * i2f(V.nl_state[6]) + 1e-3f) * 1.0002f + 0.01f;
}

template <class T>
__device__ __forceinline__ void w_cell_of(const View<T> &V, int i, float cs, int &cx, int &cy, int &cz)
{
  const double *a = XA(V, K_CA, i);
  cx = (int)floorf((float)a[0] / cs);
  cy = (int)floorf((float)a[1] / cs);
  cz = (int)floorf((float)a[2] / cs);
}

// cell coordinates packed 10 bits each (cells are >= 10 A, so +-512 cells
// spans kilometres)
__device__ __forceinline__ int w_cell_pack(int cx, int cy, int cz)
{
  return ((cx + 512) & 1023) | (((cy + 512) & 1023) << 10) | (((cz + 512) & 1023) << 20);
}

template <class T>
__device__ __forceinline__ bool w_cell_build(const BatchDev &B, const ReplicaDesc &R, int s, View<T> &V)
{
  if (!w_active(R, s) || !R.cells) return false;
  V = w_view<T>(B, R);
  return nl_needs_build(V);
}

template <class T>
__global__ void __launch_bounds__(256) k_w_cell_zero(BatchDev B, int s)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  View<T> V;
  if (!w_cell_build(B, R, s, V)) return;
  for (int k = W_TID; k < R.cell_h; k += W_NT) R.cells[k] = 0;
}

template <class T>
__global__ void __launch_bounds__(256) k_w_cell_count(BatchDev B, int s)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  View<T> V;
  if (!w_cell_build(B, R, s, V)) return;
  int h = R.cell_h, n = R.n;
  int *counts = R.cells, *bucket = R.cells + 2 * h + n, *slot = bucket + n, *cellc = slot + n;
  float cs = w_cell_size(V);
  for (int i = W_TID; i < n; i += W_NT) {
    int cx, cy, cz;
    w_cell_of(V, i, cs, cx, cy, cz);
    int b = w_cell_bucket(cx, cy, cz, h);
    bucket[i] = b;
    cellc[i] = w_cell_pack(cx, cy, cz);
    slot[i] = atomicAdd(&counts[b], 1);
  }
}

// exclusive prefix sum of the bucket counts (one workgroup per replica)
template <class T>
__global__ void __launch_bounds__(256) k_w_cell_scan(BatchDev B, int s)
{
  __shared__ int part[256];
  const ReplicaDesc &R = B.rep[blockIdx.y];
  View<T> V;
  if (!w_cell_build(B, R, s, V)) return;
  int h = R.cell_h;
  const int *counts = R.cells;
  int *offs = R.cells + h;
  int seg = (h + 255) / 256, b0 = threadIdx.x * seg, b1 = min(h, b0 + seg);
  int sum = 0;
  for (int b = b0; b < b1; ++b) sum += counts[b];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int t = 0; t < 256; ++t) { int v = part[t]; part[t] = acc; acc += v; }
  }
  __syncthreads();
  int acc = part[threadIdx.x];
  for (int b = b0; b < b1; ++b) { offs[b] = acc; acc += counts[b]; }
}

template <class T>
__global__ void __launch_bounds__(256) k_w_cell_fill(BatchDev B, int s)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  View<T> V;
  if (!w_cell_build(B, R, s, V)) return;
  int h = R.cell_h, n = R.n;
  const int *offs = R.cells + h, *bucket = R.cells + 2 * h + n, *slot = bucket + n;
  int *members = R.cells + 2 * h;
  for (int i = W_TID; i < n; i += W_NT) members[offs[bucket[i]] + slot[i]] = i;
}

// one work item per (residue, neighbouring cell): 27 n items spread over the grid
template <class T>
__global__ void __launch_bounds__(256) k_w_cell_pairs(BatchDev B, int s)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  View<T> V;
  if (!w_cell_build(B, R, s, V)) return;
  int h = R.cell_h, n = R.n;
  const int *counts = R.cells, *offs = R.cells + h, *members = R.cells + 2 * h;
  const int *cellc = R.cells + 2 * h + 3 * n;
  float reach = (float)V.P->pair_reach + (float)V.P->nl_skin;
  for (int k = W_TID; k < 27 * n; k += W_NT) {
    int i = k / 27, d = k - 27 * i;
    int c = cellc[i];
    int cx = (c & 1023) - 512 + d % 3 - 1, cy = ((c >> 10) & 1023) - 512 + (d / 3) % 3 - 1,
        cz = ((c >> 20) & 1023) - 512 + d / 9 - 1;
    int target = w_cell_pack(cx, cy, cz);
    int b = w_cell_bucket(cx, cy, cz, h);
    int m0 = offs[b], m1 = m0 + counts[b];
    for (int m = m0; m < m1; ++m) {
      int j = members[m];
      if (j <= i || cellc[j] != target) continue;
      if (near_pair(V, i, j, reach)) {
        int q = aadd_i(&V.nl_state[1], 1);
        if (q < V.nl_cap) V.nl[AW_NL_HDR + q] = (i << 16) | j;
      }
    }
  }
  for (int i = W_TID; i < n; i += W_NT) {
    const double *ca = XA(V, K_CA, i);
    V.nlx[3 * i] = ca[0]; V.nlx[3 * i + 1] = ca[1]; V.nlx[3 * i + 2] = ca[2];
    V.nlx[3 * n + i] = (double)V.ext()[i];
  }
}

template <class T>
__global__ void __launch_bounds__(64) k_w_nl_publish(BatchDev B, int s)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  if (!w_active(R, s) || threadIdx.x != 0) return;
  View<T> V = w_view<T>(B, R);
  if (V.nl) nl_publish(V, nl_needs_build(V));
}

// a phase body with per-block energy accumulators added to the replica's slots
#define W_KERNEL(NAME, BODY)                                                           \
  template <class T>                                                                   \
  __global__ void __launch_bounds__(256) NAME(BatchDev B, int s)                       \
  {                                                                                    \
    __shared__ double ered[N_ERED];                                                    \
    const ReplicaDesc &R = B.rep[blockIdx.y];                                          \
    if (!w_active(R, s)) return;                                                       \
    View<T> V = w_view<T>(B, R);                                                       \
    V.ered = ered;                                                                     \
    if (threadIdx.x < N_ERED) ered[threadIdx.x] = 0.0;                                 \
    __syncthreads();                                                                   \
    BODY;                                                                              \
    __syncthreads();                                                                   \
    energy_add(ered, R.energy);                                                        \
  }

// force sink of the wide fragment-memory kernel: a block's CA/CB forces for
// residues [rlo, rlo + span) in LDS, flushed to the replica once
struct FmSinkLds {
  double *fl;
  int rlo, span;
  template <class R>
  __device__ void operator()(int kind, int i, R fx, R fy, R fz) const
  {
    double *p = fl + ((size_t)kind * span + (i - rlo)) * 3;
    aadd_as<true>(p + 0, (double)fx);
    aadd_as<true>(p + 1, (double)fy);
    aadd_as<true>(p + 2, (double)fz);
  }
};

#define W_FM_SPAN 640  // residues of the block-local force window (3 kinds: 45 KB of LDS)

// fragment memory and the local terms: each block takes a contiguous range
// of the replica's fragment-memory items (ordered by window) and the local
// terms of a contiguous range of residues, so its forces fall on a short
// residue window [wlo, wlo + span) of all three atom kinds, accumulated in
// LDS and added into the replica's forces once per block (a global FP64
// atomic is performed at the memory side; the local terms alone issue 18
// per residue).  A block whose window would exceed W_FM_SPAN adds directly.
template <class T>
__global__ void __launch_bounds__(256) k_w_B1(BatchDev B, int s)
{
  __shared__ double ered[N_ERED];
  __shared__ double fl[3 * W_FM_SPAN * 3];
  __shared__ int lim[2];
  const ReplicaDesc &R = B.rep[blockIdx.y];
  if (!w_active(R, s)) return;
  View<T> V = w_view<T>(B, R);
  V.ered = ered;
  int n = V.n;
  // the block's residues for the local terms, whose forces reach one
  // residue either side
  int o0 = (int)((long long)n * blockIdx.x / gridDim.x), o1 = (int)((long long)n * (blockIdx.x + 1) / gridDim.x);
  if (threadIdx.x < N_ERED) ered[threadIdx.x] = 0.0;
  if (threadIdx.x == 0) {
    lim[0] = o0 < o1 ? (o0 > 0 ? o0 - 1 : 0) : 1 << 30;
    lim[1] = o0 < o1 ? (o1 < n ? o1 : n - 1) : -1;
  }
  __syncthreads();
  EAcc E;
  E.clear();
  ParamPtr P = V.P;
  int nfm = (P->fm_flag && ON(P, AW_ET_FRAGMEM)) ? R.fm_items : 0;
  typedef typename T4<T>::item ItemT;
  const ItemT *items = (const ItemT *)R.items;
  int chunk = (nfm + gridDim.x - 1) / gridDim.x;
  int k0 = blockIdx.x * chunk, k1 = min(nfm, k0 + chunk);
  for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    int ij = items[k].ij;
    atomicMin(&lim[0], ij >> 16);
    atomicMax(&lim[1], ij & 0xffff);
  }
  __syncthreads();
  int wlo = lim[0], span = lim[1] - lim[0] + 1;
  bool local = span > 0 && span <= W_FM_SPAN;
  if (local)
    for (int k = threadIdx.x; k < 9 * span; k += blockDim.x) fl[k] = 0.0;
  __syncthreads();
  // energies leave a launch from the replica's last evaluation only
  bool ef = s < 0 || s == R.nsteps - 1;
  // table mode: its pairs add into the replica's forces directly
  if (P->fm_tb_flag && ON(P, AW_ET_FRAGMEM) && R.tb_items > 0)
    E.fm += fm_table_run<T>(V.x(), V.n, V.fb, V.fs, P, V.tb_item, V.tb_curve, V.fault, W_TID, R.tb_items, W_NT);
  int kk = k0 + (int)threadIdx.x, ks = (int)blockDim.x;
  // inlined, unpipelined, first batches of 8 memories: a lane has a few
  // items, and the kernel keeps more waves per SIMD with the smaller
  // register footprint
  const int W_FMB = sizeof(T) == 4 ? 8 : 4;
  if (local) {
    typedef FmSinkLds Sk;
    Sk sink{fl, wlo, span};
    if (ef) E.fm += fm_run_body<T, false, true, Sk, false, W_FMB>(V.x(), V.n, V.rf, V.coef, items, kk, k1, ks, sink);
    else fm_run_body<T, false, false, Sk, false, W_FMB>(V.x(), V.n, V.rf, V.coef, items, kk, k1, ks, sink);
    // the local terms add through the view into the same window: atom
    // (kind, i) at fl + 3 (kind span + i - wlo), i.e. a force base 3 wlo
    // doubles before the window (formed as an integer: the base itself lies
    // outside it, every address used lies inside)
    V.fb = (double *)((uintptr_t)fl - (uintptr_t)wlo * 3 * sizeof(double));
    V.fs = span;
  } else {
    typedef FmSinkForce<false> Sk;
    Sk sink{aw_cast<false>(V.f()), V.n};
    if (ef) E.fm += fm_run_body<T, false, true, Sk, false, W_FMB>(V.x(), V.n, V.rf, V.coef, items, kk, k1, ks, sink);
    else fm_run_body<T, false, false, Sk, false, W_FMB>(V.x(), V.n, V.rf, V.coef, items, kk, k1, ks, sink);
  }
  for (int k = o0 + (int)threadIdx.x; k < o1; k += blockDim.x) residue_local(V, k, E);
  __syncthreads();
  if (local)
    for (int k = threadIdx.x; k < 9 * span; k += blockDim.x) {
      double v = fl[k];
      if (v != 0.0) {
        int kind = k / (3 * span), rem = k - kind * 3 * span;
        __hip_atomic_fetch_add(V.f() + ((size_t)kind * n + wlo) * 3 + rem, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  eflush(E, V.ered);
  __syncthreads();
  energy_add(ered, R.energy);
}

// the coeff-toggled terms (awsem_extra.h) over each replica's lanes, after
// the list is published (the barrier sweeps it)
template <class T>
__global__ void __launch_bounds__(256) k_w_X(BatchDev B, int s)
{
  __shared__ double ered[N_ERED];
  const ReplicaDesc &R = B.rep[blockIdx.y];
  if (!w_active(R, s)) return;
  View<T> V = w_view<T>(B, R);
  V.ered = ered;
  if (threadIdx.x < N_ERED) ered[threadIdx.x] = 0.0;
  __syncthreads();
  xt_pass(V, R, W_TID, W_NT, s < 0 || s == R.nsteps - 1);
  __syncthreads();
  energy_add(ered, R.energy);
}

// the list scan: densities in place, DSSP / P_AP / excluded-volume pairs
// into their queues and the density pairs for D and F; without a valid
// list every pair is visited and evaluated in place
template <class T>
__device__ __forceinline__ void w_scan(const View<T> &V, bool listless_only)
{
  EAcc E;
  E.clear();
  if (V.wq && V.nl_state[7] > 0) {
    if (listless_only) return;  // the windowed scan takes this replica
    int cap = V.nl_cap;
    AW_G int *qd = V.wq, *qp = V.wq + 2 * cap, *qe = V.wq + 3 * cap, *qs = V.wq + 4 * cap;
    for_pairs(V, W_TID, W_NT, [&](int i, int j) { pair_B<T, true>(V, i, j, E, qd, qp, qe, qs); });
  } else {
    for_pairs(V, W_TID, W_NT, [&](int i, int j) { pair_B(V, i, j, E); });
  }
  eflush(E, V.ered);
}
// without windows, every replica; with them (lo = 1), the replicas whose
// list is not valid (every pair evaluated in place)
template <class T>
__global__ void __launch_bounds__(256) k_w_B2(BatchDev B, int s, int lo)
{
  __shared__ double ered[N_ERED];
  const ReplicaDesc &R = B.rep[blockIdx.y];
  if (!w_active(R, s)) return;
  View<T> V = w_view<T>(B, R);
  V.ered = ered;
  if (threadIdx.x < N_ERED) ered[threadIdx.x] = 0.0;
  __syncthreads();
  w_scan(V, lo != 0);
  __syncthreads();
  energy_add(ered, R.energy);
}

// the queue passes, one launch per queue: DSSP (q = 0), P_AP (1),
// excluded volume (2)
template <class T>
__device__ __forceinline__ void w_queue(const View<T> &V, int q)
{
  EAcc E;
  E.clear();
  if (V.wq && V.nl_state[7] > 0) {
    int cap = V.nl_cap;
    if (q == 0)
      for_packed((const AW_G int *)V.wq, 0, V.nl_state[2], W_TID, W_NT, [&](int i, int j) { dssp_term(V, i, j, E); });
    else if (q == 1)
      for_packed((const AW_G int *)V.wq + 2 * cap, 0, V.nl_state[3], W_TID, W_NT,
                 [&](int i, int j) { pap_term(V, i, j, E); });
    else
      for_packed((const AW_G int *)V.wq + 3 * cap, 0, V.nl_state[4], W_TID, W_NT,
                 [&](int i, int j) { ev_pair(V, i, j, V.rinfo[i], V.rinfo[j], E); });
  }
  eflush(E, V.ered);
}
W_KERNEL(k_w_B3d, w_queue(V, 0))
W_KERNEL(k_w_B3p, w_queue(V, 1))
W_KERNEL(k_w_B3e, w_queue(V, 2))
W_KERNEL(k_w_C, phase_C(V, W_TID, W_NT))
W_KERNEL(k_w_D, phase_D(V, W_TID, W_NT))
W_KERNEL(k_w_F, phase_F(V, W_TID, W_NT))

// ---- force windows.  A global FP64 atomic add is performed at the memory
// side (device-coherent across the XCDs' L2 caches), at about 2e10 adds/s
// for the whole GPU; the DSSP, P_AP and density-pair force passes issue
// about 36, 14 and 20 of them per item, which bounds them.  Their windowed
// form gives each workgroup (1024 threads, a few per replica) an LDS copy
// of the two atom kinds the pass writes, adds into it with LDS atomics and
// adds the window into the replica's forces once: q = 0, the DSSP and P_AP
// queues (CA and O forces only; O rows start at 2 fs with fs = ceil(n / 2),
// so the two kinds fill 6 fs + 3 n doubles); q = 1, the density-pair forces
// of water, burial and helix (CA and CB: site atoms only, 6 n doubles);
// q = 2, the list scan's densities rho_w and rho_h (2 n doubles); q = 3,
// phase D: helix forces (CA and O, as q = 0) and water xi (n doubles after
// them).
#define W_WIN_NT 1024

__host__ __device__ inline int w_win_fs(int n, int q) { return (q == 0 || q == 3) ? (n + 1) / 2 : n; }
__host__ __device__ inline int w_win_doubles(int n, int q)
{
  return q == 0 ? 6 * w_win_fs(n, 0) + 3 * n : q == 1 ? 6 * n : q == 2 ? 2 * n : 6 * w_win_fs(n, 3) + 4 * n;
}

template <class T>
__global__ void __launch_bounds__(W_WIN_NT) k_w_win(BatchDev B, int s, int q)
{
  extern __shared__ double win[];
  __shared__ double ered[N_ERED];
  const ReplicaDesc &R = B.rep[blockIdx.y];
  if (!w_active(R, s)) return;
  View<T> V = w_view<T>(B, R);
  int n = R.n;
  // a replica without a valid list: its scan is the direct kernel's
  // (k_w_B2, launched beside this one, takes exactly those replicas)
  if (q == 2 && !(V.wq && V.nl_state[7] > 0)) return;
  int fs = w_win_fs(n, q), wn = w_win_doubles(n, q);
  for (int k = threadIdx.x; k < wn; k += blockDim.x) win[k] = 0.0;
  if (threadIdx.x < N_ERED) ered[threadIdx.x] = 0.0;
  __syncthreads();
  V.ered = ered;
  AW_G double *g = (AW_G double *)V.f();  // the replica's force array [3][n][3]
  V.fb = win;
  V.fs = fs;
  if (q == 0) {
    EAcc E;
    E.clear();
    if (V.wq && V.nl_state[7] > 0) {
      int cap = V.nl_cap;
      for_packed((const AW_G int *)V.wq, 0, V.nl_state[2], W_TID, W_NT, [&](int i, int j) { dssp_term(V, i, j, E); });
      for_packed((const AW_G int *)V.wq + 2 * cap, 0, V.nl_state[3], W_TID, W_NT,
                 [&](int i, int j) { pap_term(V, i, j, E); });
    }
    eflush(E, V.ered);
  } else if (q == 1) {
    phase_F(V, W_TID, W_NT);
  } else if (q == 2) {
    // the launcher takes this form only for a valid list (queued scan: the
    // densities are its only in-place sums)
    EAcc E;
    E.clear();
    int cap = V.nl_cap;
    AW_G int *qd = V.wq, *qp = V.wq + 2 * cap, *qe = V.wq + 3 * cap, *qs = V.wq + 4 * cap;
    DensWin dw{win};
    for_pairs(V, W_TID, W_NT,
              [&](int i, int j) { pair_B<T, true, DensWin>(V, i, j, E, qd, qp, qe, qs, nullptr, dw); });
    eflush(E, V.ered);
  } else {
    phase_D(V, W_TID, W_NT, true, XiWin{win + 6 * fs + 3 * n});
  }
  __syncthreads();
  if (q == 2) {
    // rho_w, rho_h into the slab's per-residue scalars
    AW_G double *sc = (AW_G double *)(V.s + 27LL * n);
    for (int k = threadIdx.x; k < 2 * n; k += blockDim.x) {
      double v = win[k];
      if (v != 0.0) __hip_atomic_fetch_add(sc + AW_SC_STRIDE * (k >> 1) + (k & 1), v, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
    }
  } else {
    // the window into the replica's forces: CA rows, then O (q = 0, 3) or
    // CB rows (q = 1)
    bool ca_o = q != 1;
    int o2 = ca_o ? 6 * fs : 3 * n, k2 = ca_o ? 6 * n : 3 * n;
    for (int k = threadIdx.x; k < 6 * n; k += blockDim.x) {
      double v = k < 3 * n ? win[k] : win[o2 + k - 3 * n];
      int gk = k < 3 * n ? k : k2 + k - 3 * n;
      if (v != 0.0) __hip_atomic_fetch_add(g + gk, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (q == 3) {
      // water xi into the slab's per-residue scalars
      AW_G double *xi = (AW_G double *)(V.s + 27LL * n + 6);
      const double *xw = win + 6 * fs + 3 * n;
      for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double v = xw[k];
        if (v != 0.0) __hip_atomic_fetch_add(xi + AW_SC_STRIDE * k, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
  energy_add(ered, R.energy);
}

template <class T>
static void prepare_wide_t()
{
  int dev = 0, lim = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lim <= 0) lim = 65536;
  (void)hipFuncSetAttribute((const void *)k_w_win<T>, hipFuncAttributeMaxDynamicSharedMemorySize, lim - 512);
}

template <class T>
__global__ void __launch_bounds__(256) k_w_int1(BatchDev B, LgParams lg, int s)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  if (!w_active(R, s)) return;
  View<T> V = w_view<T>(B, R);
  integrate1(V, (AW_G double *)R.v, lg, W_TID, W_NT);
}

template <class T>
__global__ void __launch_bounds__(256) k_w_int2(BatchDev B, LgParams lg, int s)
{
  const ReplicaDesc &R = B.rep[blockIdx.y];
  if (!w_active(R, s)) return;
  View<T> V = w_view<T>(B, R);
  integrate2(V, (AW_G double *)R.v, lg, R, s, W_TID, W_NT);
}

// the windowed passes' launch: wg workgroups per replica, lds[q] bytes of
// window for pass q (0: the direct kernel, for replicas too large for it)
struct WWin {
  int wg;
  size_t lds[4];
  int xt;  // the coeff-toggled terms' pass (k_w_X)
};

template <class T>
static void w_eval(const BatchDev &B, dim3 g, hipStream_t st, int s, const WWin &w)
{
  dim3 g1(1, g.y);
  hipLaunchKernelGGL(k_w_A<T>, g, dim3(256), 0, st, B, s);
  hipLaunchKernelGGL(k_w_nl_build<T>, g, dim3(256), 0, st, B, s);
  hipLaunchKernelGGL(k_w_cell_zero<T>, g, dim3(256), 0, st, B, s);
  hipLaunchKernelGGL(k_w_cell_count<T>, g, dim3(256), 0, st, B, s);
  hipLaunchKernelGGL(k_w_cell_scan<T>, g1, dim3(256), 0, st, B, s);
  hipLaunchKernelGGL(k_w_cell_fill<T>, g, dim3(256), 0, st, B, s);
  hipLaunchKernelGGL(k_w_cell_pairs<T>, g, dim3(256), 0, st, B, s);
  hipLaunchKernelGGL(k_w_nl_publish<T>, g1, dim3(64), 0, st, B, s);
  if (w.xt) hipLaunchKernelGGL(k_w_X<T>, g, dim3(256), 0, st, B, s);
  hipLaunchKernelGGL(k_w_B1<T>, g, dim3(256), 0, st, B, s);
  dim3 gw(w.wg, g.y);
  hipLaunchKernelGGL(k_w_B2<T>, g, dim3(256), 0, st, B, s, w.lds[2] ? 1 : 0);
  if (w.lds[2]) hipLaunchKernelGGL(k_w_win<T>, gw, dim3(W_WIN_NT), w.lds[2], st, B, s, 2);
  if (w.lds[0]) {
    hipLaunchKernelGGL(k_w_win<T>, gw, dim3(W_WIN_NT), w.lds[0], st, B, s, 0);
  } else {
    hipLaunchKernelGGL(k_w_B3d<T>, g, dim3(256), 0, st, B, s);
    hipLaunchKernelGGL(k_w_B3p<T>, g, dim3(256), 0, st, B, s);
  }
  hipLaunchKernelGGL(k_w_B3e<T>, g, dim3(256), 0, st, B, s);
  hipLaunchKernelGGL(k_w_C<T>, g, dim3(256), 0, st, B, s);
  if (w.lds[3])
    hipLaunchKernelGGL(k_w_win<T>, gw, dim3(W_WIN_NT), w.lds[3], st, B, s, 3);
  else
    hipLaunchKernelGGL(k_w_D<T>, g, dim3(256), 0, st, B, s);
  if (w.lds[1])
    hipLaunchKernelGGL(k_w_win<T>, gw, dim3(W_WIN_NT), w.lds[1], st, B, s, 1);
  else
    hipLaunchKernelGGL(k_w_F<T>, g, dim3(256), 0, st, B, s);
}

// K steps of every replica (mode 1), or one force evaluation (mode 0);
// run.grid = workgroups per replica (0: enough for about 2048 in total)
template <class T>
static int launch_wide_t(const AwBatchPtrs &p, const AwRun &run, int max_steps, hipStream_t st)
{
  if (p.nrep <= 0) return 0;
  BatchDev B = to_dev(p);
  LgParams lg = lg_make(run.dt, run.damp);
  int bx = run.grid > 0 ? run.grid : (2048 + p.nrep - 1) / p.nrep;
  bx = bx < 1 ? 1 : (bx > 64 ? 64 : bx);
  dim3 g(bx, p.nrep);
  // windowed passes: about one 1024-thread workgroup per CU in total, at
  // most 16 per replica (each adds its whole window once); replicas whose
  // window exceeds the LDS take the direct passes (AWSEM_WIDE_WINDOW=0:
  // always)
  WWin w;
  w.xt = p.extras;
  w.wg = (256 + p.nrep - 1) / p.nrep;
  w.wg = w.wg < 1 ? 1 : (w.wg > 16 ? 16 : w.wg);
  bool win_on = true;
  if (const char *e = getenv("AWSEM_WIDE_WINDOW")) win_on = atoi(e) != 0;
  for (int q = 0; q < 4; ++q) {
    w.lds[q] = (size_t)w_win_doubles(p.max_n, q) * sizeof(double);
    if (!win_on || w.lds[q] > (size_t)aw_lds_limit_bytes() - 512) w.lds[q] = 0;
  }
  hipLaunchKernelGGL(k_w_begin<T>, g, dim3(256), 0, st, B, run.mode);
  if (run.mode == 1) {
    w_eval<T>(B, g, st, -1, w);
    for (int s = 0; s < max_steps; ++s) {
      hipLaunchKernelGGL(k_w_int1<T>, g, dim3(256), 0, st, B, lg, s);
      w_eval<T>(B, g, st, s, w);
      hipLaunchKernelGGL(k_w_int2<T>, g, dim3(256), 0, st, B, lg, s);
    }
  } else {
    w_eval<T>(B, g, st, -2, w);
  }
  hipLaunchKernelGGL(k_w_end<T>, g, dim3(256), 0, st, B, run.mode);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
