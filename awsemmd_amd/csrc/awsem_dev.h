// awsem_dev.h -- device-side batch description, replica views and the
// resident engine, shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "awsem_terms.h"
#include "awsem_extra.h"
#include "awsem_launch.h"

// ------------------------------------------------------------------------
// energy epilogue: ered[N_ERED] (LDS) -> out[AW_N_ET]

// phase engine: blocks add their partial sums
__device__ __forceinline__ void energy_add(const double *ered, double *out)
{
  if (threadIdx.x < N_ERED) {
    double v = ered[threadIdx.x];
    if (v != 0.0) atomicAdd(&out[eacc_slot[threadIdx.x]], v);
  }
}

// resident engine: every slot written, slot 0 the total
__device__ __forceinline__ void energy_write(const double *ered, double *out)
{
  int t = threadIdx.x;
  if (t < AW_N_ET) {
    double v = 0.0;
    if (t == 0) {
      for (int k = 0; k < N_ERED; ++k) v += ered[k];
    } else {
      for (int k = 0; k < N_ERED; ++k)
        if (eacc_slot[k] == t) v = ered[k];
    }
    out[t] = v;
  }
}

// ------------------------------------------------------------------------
// batch description and views

struct BatchDev {
  const ReplicaDesc *rep;     // [nrep]
  int nrep;
  const FFParams *P;
  int *queue;                 // work-queue counter (resident engine)
  const int *order;           // replica order (longest first)
  unsigned long long *prof;   // optional phase clock: [2][16] (first replica, all replicas)
  // fragment-memory sharing: the replicas other workgroups may help
  int nshare;
  const int *share_rep, *share_n;       // [nshare] replica, chunks
  unsigned long long *share_word;       // [nshare] claim words
  int share_spread;                     // a helper's first look: the (blockIdx mod spread)-th open replica (power of 2)
};

// bytes of a workgroup's LDS left to the resident kernel's static arrays
#define AW_LDS_STATIC_RESERVE 256

// doubles of per-replica working storage (x, xv, f, the per-residue
// scalars, ext, rinfo copy)
__host__ __device__ inline long long slab_doubles(int n) { return 27LL * n + (long long)AW_SC_STRIDE * n + 2LL * ((n + 1) / 2) + 8; }

template <class T>
__device__ __forceinline__ View<T> make_view(const BatchDev &B, const ReplicaDesc &R, double *slab)
{
  View<T> V;
  V.n = R.n; V.nch = R.nch;
  V.rinfo = R.rinfo;
  V.aps = (const AW_G double *)R.aps;
  V.ch_start = (const AW_G int *)R.ch_start;
  V.ch_len = (const AW_G int *)R.ch_len;
  V.s = slab;
  V.fb = slab + 18LL * R.n;
  V.fs = R.n;
  V.ered = nullptr;
  V.P = (ParamPtr)B.P;
  V.fm_items = R.fm_items;
  V.fm_item_base = 0;
  V.tb_items = R.tb_items;
  V.tb_item = (const AW_G int *)R.tb_item;
  V.tb_curve = R.tb_curve;
  V.fault = (AW_G int *)R.fault;
  V.coef = (const T *)R.coef;
  V.rf = R.rf;
  V.items = R.items;
  V.nl = nullptr;
  V.nlx = nullptr;
  V.nl_cap = 0;
  V.nl_state = nullptr;
  V.wq = nullptr;
  V.sh_n = 0;
  V.sh_word = nullptr;
  V.sh_chunks = nullptr;
  V.sh_mask = nullptr;
  return V;
}

static inline BatchDev to_dev(const AwBatchPtrs &p)
{
  BatchDev B;
  B.rep = (const ReplicaDesc *)p.rep;
  B.nrep = p.nrep;
  B.P = (const FFParams *)p.params;
  B.queue = p.queue;
  B.order = p.order;
  B.prof = p.prof;
  B.nshare = p.nshare;
  B.share_rep = p.share_rep;
  B.share_n = p.share_n;
  B.share_word = p.share_word;
  B.share_spread = p.share_spread;
  return B;
}

// ------------------------------------------------------------------------
// resident engine: one workgroup per replica, nsteps inside the launch

__device__ __forceinline__ void prof_mark(unsigned long long *prof, bool first, int k, unsigned long long &t)
{
  if (prof && threadIdx.x == 0) {
    unsigned long long now = wall_clock64();
    atomicAdd(&prof[16 + k], now - t);
    if (first) atomicAdd(&prof[k], now - t);
    t = now;
  }
}

// sub-phase clock of the queued pair pass
struct PairMarks {
  unsigned long long *prof;
  bool first;
  unsigned long long t;
  __device__ __forceinline__ void mark(int k) { prof_mark(prof, first, k, t); }
};

// ------------------------------------------------------------------------
// fragment-memory sharing (resident engine, replica slab in LDS)
//
// Every access another workgroup reads or writes (the position copy, the
// accumulator, the claim words and done counters) is a device-scope
// atomic: loads and stores carry the device coherence bit (they miss
// stale cache lines and write through), read-modify-writes are performed
// device-coherently.  Order between them is the hardware's: a wave waits
// for its outstanding memory operations (vm_drain) before it publishes a
// claim word or a done count, and reads the data only after it has read
// the word or count.  No cache-wide write-back or invalidation is needed,
// since nothing is shared through plain (cached) accesses.

// wait until this wave's global memory operations (stores and atomics
// included) have completed; a compiler barrier on both sides
__device__ __forceinline__ void vm_drain()
{
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xF70);  // vmcnt(0), the other counters untouched
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
template <class V>
__device__ __forceinline__ V ld_dev(const V *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
template <class V>
__device__ __forceinline__ void st_dev(V *p, V v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// spin limits (a spin is one device load and a short sleep, about a
// microsecond): an owner that waits longer than about 0.1 s for its chunks
// gives up, raises the batch's error flag and evaluates its fragment
// memory itself for the rest of the launch; a helper idle for about 0.3 s
// leaves
#define AW_SHARE_WAIT_MAX (1 << 17)
#define AW_SHARE_IDLE_MAX (1 << 18)

// the owner, before its evaluation's passes: the CA and CB positions to
// the replica's copy; every thread drains its stores before the barrier
// that ends phase A
__device__ __forceinline__ void share_snapshot(const ReplicaDesc &R, const double *x, int n)
{
  for (int k = threadIdx.x; k < 6 * n; k += blockDim.x) st_dev(R.sh_x + k, x[k]);
}

// the owner, after that barrier: open the chunks of this evaluation
__device__ __forceinline__ void share_publish(const ReplicaDesc &R, bool ef)
{
  if (threadIdx.x == 0) {
    st_dev(R.sh_done, 0);
    vm_drain();
    __hip_atomic_exchange(R.sh_word, ef ? (1ull << 32) : 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// the owner, after its evaluation (every wave has left its claim loop, so
// every chunk is claimed): wait for the chunks taken elsewhere, add their
// forces (and energy) and clear the accumulator for the next evaluation.
// Returns false (in every thread) when the wait gave up.
template <class T>
__device__ __forceinline__ bool share_finish(const View<T> &V, const ReplicaDesc &R, bool ef, int *err, int *s_ok)
{
  unsigned long long mine = *V.sh_mask;
  int away = V.sh_n - __popcll(mine);
  if (away == 0) return true;  // uniform: the mask is read after a barrier
  if (threadIdx.x == 0) {
    // chunks evaluated elsewhere, summed over the batch's launches
    __hip_atomic_fetch_add(err + 16, away, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int spins = 0, ok = 1;
    while (ld_dev(R.sh_done) < away) {
      if (++spins > AW_SHARE_WAIT_MAX) {
        st_dev(err, 1);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *s_ok = ok;
  }
  __syncthreads();
  if (!*s_ok) return false;
  double *acc = R.sh_acc;
  double *f = V.f();
  int m = 6 * V.n;
  for (int k = threadIdx.x; k < m; k += blockDim.x) {
    double a = ld_dev(acc + k);
    if (a != 0.0) {
      f[k] += a;
      st_dev(acc + k, 0.0);
    }
  }
  if (threadIdx.x == 0) {
    double a = ld_dev(acc + m);
    if (ef) V.ered[9] += a;  // the fragment-memory accumulator (eacc_slot)
    st_dev(acc + m, 0.0);
  }
  __syncthreads();
  return true;
}

// a workgroup with no replica left helps the shared ones until every
// replica of the launch is finished: wave 0 looks for a replica with
// unclaimed chunks and claims AW_SHARE_TAKE of them, one per wave; a wave
// copies its chunk's residue span into its own LDS region, evaluates the
// chunk's items there and adds the span's forces into the replica's
// accumulator before the chunk counts as done
template <class T>
__device__ void share_help(const BatchDev &B, double *smem)
{
  typedef typename T4<T>::item ItemT;
  typedef FmSinkForce<true> Sk;
  __shared__ int s_q, s_c0, s_ef, s_quit;
  const int SP = AW_SHARE_SPAN;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double *xw = smem + wave * 12 * SP;  // CA, CB rows of the span
  double *fw = xw + 6 * SP;            // their forces
  int idle = 0;
  for (;;) {
    __syncthreads();  // the previous round's broadcast has been read
    if (wave == 0) {
      int pick = -1;
      // the first open replica, longest first, from the (blockIdx mod
      // spread)-th on (spread 1: strictly the longest open one)
      int rot = blockIdx.x & (B.share_spread - 1);
      for (int b = 0; b < B.nshare && pick < 0; b += 64) {
        int q = b + ((lane + rot) & 63);
        bool open = false;
        if (q < B.nshare) open = (unsigned)ld_dev(B.share_word + q) < (unsigned)B.share_n[q];
        unsigned long long m = __ballot(open);
        if (m) {
          int l = __ffsll((long long)m) - 1;
          pick = b + ((l + rot) & 63);
        }
      }
      int c0 = 0, ef = 0, quit = 0;
      if (pick >= 0) {
        if (lane == 0) {
          unsigned long long w = __hip_atomic_fetch_add(B.share_word + pick, (unsigned long long)AW_SHARE_TAKE,
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          c0 = (int)(unsigned)w;
          ef = (int)(w >> 32);
        }
        c0 = __builtin_amdgcn_readfirstlane(c0);
        ef = __builtin_amdgcn_readfirstlane(ef);
        if (c0 >= B.share_n[pick]) pick = -1;  // taken by others meanwhile: counts as an idle look
        else idle = 0;
      }
      if (pick < 0) {
        int fin = 0;
        if (lane == 0) fin = ld_dev(B.queue + 16);
        fin = __builtin_amdgcn_readfirstlane(fin);
        quit = fin >= B.nrep || ++idle > AW_SHARE_IDLE_MAX;
      }
      if (lane == 0) { s_q = pick; s_c0 = c0; s_ef = ef; s_quit = quit; }
    }
    __syncthreads();
    int q = s_q, c0 = s_c0;
    bool ef = s_ef != 0;
    if (s_quit) break;
    if (q < 0) {
      __builtin_amdgcn_s_sleep(8);
      continue;
    }
    const ReplicaDesc &R = B.rep[B.share_rep[q]];
    int c = c0 + wave;
    if (wave >= AW_SHARE_TAKE || c >= R.sh_n) continue;
    int n = R.n;
    ShareChunk ch = R.sh_chunks[c];
    int lo = ch.lo;
    // the span's rows (rows past the replica's end stay zero)
    for (int k = lane; k < 6 * SP; k += 64) {
      int kind = k / (3 * SP), row = (k - kind * 3 * SP) / 3, comp = k - kind * 3 * SP - 3 * row;
      int res = lo + row;
      xw[k] = res < n ? ld_dev(R.sh_x + (kind * n + res) * 3 + comp) : 0.0;
      fw[k] = 0.0;
    }
    // LDS writes of other lanes of this wave are complete before the reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const ItemT *items = (const ItemT *)R.items;
    Sk sink{aw_cast<true>(fw), SP};
    double e = 0.0;
    if (ef)
      e = fm_run<T, true, true, Sk, true>(xw, SP, R.rf, (const T *)R.coef, items, ch.k0 + lane, ch.k1, 64, sink, lo);
    else
      fm_run<T, true, false, Sk, true>(xw, SP, R.rf, (const T *)R.coef, items, ch.k0 + lane, ch.k1, 64, sink, lo);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    double *acc = R.sh_acc;
    for (int k = lane; k < 6 * SP; k += 64) {
      int kind = k / (3 * SP), row = (k - kind * 3 * SP) / 3, comp = k - kind * 3 * SP - 3 * row;
      int res = lo + row;
      double v = fw[k];
      if (res < n && v != 0.0)
        __hip_atomic_fetch_add(acc + (kind * n + res) * 3 + comp, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (ef) {
      e = wave_sum_d(e);
      if (lane == 0 && e != 0.0) __hip_atomic_fetch_add(acc + 6 * n, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // the chunk's additions are complete before it counts as done
    vm_drain();
    if (lane == 0) __hip_atomic_fetch_add(R.sh_done, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// VAR: bit 0 the fragment-memory table pass (table mode), bit 1 the pass of
// the coeff-toggled terms; instances 0, 1 and 3 exist, so batches without
// either run a kernel with neither call
template <class T, bool LDSM, int VAR>
__device__ __forceinline__ void force_eval_block(const View<T> &V, const ReplicaDesc &R, unsigned long long *prof,
                                                 bool first, bool ef)
{
  int t0 = threadIdx.x, ts = blockDim.x;
  unsigned long long t = prof ? wall_clock64() : 0;
  if (threadIdx.x < N_ERED) V.ered[threadIdx.x] = 0.0;
  if (V.sh_n > 0) {
    share_snapshot(R, V.x(), V.n);
    if (threadIdx.x == 0) *V.sh_mask = 0ull;
  }
  phase_A(V, t0, ts);
  if (V.sh_n > 0) vm_drain();  // the copy is complete before the chunks open
  __syncthreads();
  if (V.sh_n > 0) share_publish(R, ef);
  prof_mark(prof, first, 0, t);
  // table mode: the table pass runs here, where few registers are live (its
  // call inside the pair pass costs spills there); forces are zeroed
  if constexpr ((VAR & 1) != 0) {
    fm_table_pass(V, t0, ts, ef);
    if (prof) {
      __syncthreads();
      prof_mark(prof, first, 8, t);
    }
  }
  if (V.nl) {
    bool build = nl_needs_build(V);
    if (build) phase_nl_build(V, t0, ts);
    __syncthreads();
    if (threadIdx.x == 0) nl_publish(V, build);
    __syncthreads();
  }
  prof_mark(prof, first, 11, t);
  // coeff-toggled terms: positions and the list only, forces by atomics
  // (no barrier before phase B, which touches neither their inputs nor
  // their accumulators)
  if constexpr ((VAR & 2) != 0) xt_pass(V, R, t0, ts, ef);
  // phase B in one part, or (clocked run) its three parts separately; one
  // call site keeps a single copy of the code
  int nparts = prof ? 3 : 1;
  for (int q = 0; q < nparts; ++q) {
    PairMarks pm{prof, first, t};
    phase_B<T, true, PairMarks, LDSM, false>(V, t0, ts, prof ? (1 << q) : 7, prof ? &pm : nullptr, ef);
    __syncthreads();
    prof_mark(prof, first, prof ? 8 + q : 10, t);
  }
  phase_C(V, t0, ts, ef);
  __syncthreads();
  prof_mark(prof, first, 2, t);
  phase_D(V, t0, ts, ef);
  __syncthreads();
  prof_mark(prof, first, 3, t);
  phase_F(V, t0, ts, ef);
  __syncthreads();
  prof_mark(prof, first, 5, t);
}

// VAR: the passes compiled in (force_eval_block)
template <class T, bool LDSM, int NT, int VAR>
__global__ void __launch_bounds__(NT) k_resident(BatchDev B, const FFParams *__restrict__ Pr, LgParams lg, int mode)
{
  static_assert(NT == 64 * AW_SHARE_TAKE, "a helping workgroup evaluates one claimed chunk per wave");
  extern __shared__ double smem[];
  __shared__ int s_rep;
  __shared__ double ered[N_ERED];
  __shared__ __attribute__((aligned(8))) int nl_state[8];  // [2..5]: two 64-bit queue-counter words
  __shared__ unsigned long long sh_mask;
  __shared__ int sh_ok;
  for (;;) {
    if (threadIdx.x == 0) s_rep = atomicAdd(B.queue, 1);
    __syncthreads();
    // wave-uniform (an LDS broadcast is not known to be): the descriptor,
    // n and every slab address derived from it then live in scalar
    // registers instead of vector registers (and their spills)
    int slot = __builtin_amdgcn_readfirstlane(s_rep);
    __syncthreads();
    if (slot >= B.nrep) break;
    int r = B.order[slot];
    // fields are read where used: holding the whole descriptor in registers
    // across the step loop costs spills
    const ReplicaDesc &R = B.rep[r];
    double *slab = LDSM ? smem : R.scratch;
    View<T> V = make_view<T>(B, R, slab);
    V.P = (ParamPtr)Pr;  // read-only: invariant uniform loads through the scalar cache
    V.ered = ered;
    V.nl = R.nl_cap > 0 ? (AW_G int *)R.nl : nullptr;
    V.nlx = (AW_G double *)R.nlx;
    V.nl_cap = R.nl_cap;
    V.nl_state = nl_state;
    V.wq = V.nl ? V.nl + AW_NL_HDR + R.nl_cap : nullptr;
    // a replica shares only evaluations that compute fragment memory
    if (LDSM && B.nshare > 0 && R.sh_n > 0 && Pr->fm_flag && (Pr->term_mask & (1u << AW_ET_FRAGMEM))) {
      V.sh_n = R.sh_n;
      V.sh_word = R.sh_word;
      V.sh_chunks = R.sh_chunks;
      V.sh_mask = &sh_mask;
    }
    if (threadIdx.x < 7) nl_state[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      if ((R.flags & AW_RF_NL_RESET) && V.nl) V.nl[1] = 0;
      nl_state[7] = 0;
      if (V.nl) nl_state_init(V);
    }
    int n = R.n;
    AW_G double *gx = (AW_G double *)R.x, *gv = (AW_G double *)R.v, *gf = (AW_G double *)R.f;
    for (int k = threadIdx.x; k < 9 * n; k += blockDim.x) V.x()[k] = gx[k];
    int *ri = V.rinfo_copy();
    for (int k = threadIdx.x; k < n; k += blockDim.x) ri[k] = V.rinfo[k];
    // dynamics continuing from the previous launch: its final forces are
    // still valid, so the first step integrates with them directly
    int nst = mode == 1 ? R.nsteps : 0;
    bool have_f = mode == 1 && (R.flags & AW_RF_F_VALID) && nst > 0;
    if (have_f)
      for (int k = threadIdx.x; k < 9 * n; k += blockDim.x) V.f()[k] = gf[k];
    __syncthreads();
    V.rinfo = ri;
    bool first = slot == 0;
    // s = -1: the force evaluation before the first step (one call site)
    for (int s = have_f ? 0 : -1; s < nst; ++s) {
      unsigned long long t = B.prof ? wall_clock64() : 0;
      if (s >= 0) {
        integrate1(V, gv, lg, threadIdx.x, blockDim.x);
        __syncthreads();
        prof_mark(B.prof, first, 6, t);
      }
      // energies leave the launch from its last evaluation only
      force_eval_block<T, LDSM, VAR>(V, R, B.prof, first, s == nst - 1);
      // a replica whose wait gave up evaluates its own fragment memory from
      // then on (the launch is reported as failed)
      if (V.sh_n > 0 && !share_finish(V, R, s == nst - 1, B.queue + 32, &sh_ok)) V.sh_n = 0;
      if (s >= 0) {
        t = B.prof ? wall_clock64() : 0;
        integrate2(V, gv, lg, R, s, threadIdx.x, blockDim.x);
        __syncthreads();
        prof_mark(B.prof, first, 7, t);
      }
    }
    if (mode == 1)
      for (int k = threadIdx.x; k < 9 * n; k += blockDim.x) gx[k] = V.x()[k];
    for (int k = threadIdx.x; k < 9 * n; k += blockDim.x) gf[k] = V.f()[k];
    energy_write(ered, R.energy);
    __syncthreads();
    if (B.nshare > 0 && threadIdx.x == 0)
      __hip_atomic_fetch_add(B.queue + 16, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // no replica left: help the shared ones (the slab is free now)
  if (LDSM && B.nshare > 0) share_help<T>(B, smem);
}

// the work-queue counter and the finished-replica count of the resident
// launch (a template: this header is compiled into several objects)
template <int D>
__global__ void k_zero_queue(int *p)
{
  if (threadIdx.x == 0) { p[0] = 0; p[16] = 0; }
}

template <class T>
static void prepare_resident_t()
{
  int dev = 0, lim = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lim <= 0) lim = 65536;
  (void)hipFuncSetAttribute((const void *)k_resident<T, true, 512, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            lim - AW_LDS_STATIC_RESERVE);
  (void)hipFuncSetAttribute((const void *)k_resident<T, true, 512, 1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            lim - AW_LDS_STATIC_RESERVE);
  (void)hipFuncSetAttribute((const void *)k_resident<T, true, 512, 3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            lim - AW_LDS_STATIC_RESERVE);
}

// launches the resident engine over p.rep[0..nrep) (descriptors already on
// the device, p.queue zeroed by the launcher)
template <class T>
static int launch_resident_t(const AwBatchPtrs &p, const AwRun &run, hipStream_t st)
{
  BatchDev B = to_dev(p);
  LgParams lg = lg_make(run.dt, run.damp);
  int mode = run.mode;
  if (p.nrep <= 0) return 0;
  // a kernel, not a memset node, so that the sequence can be graph-captured
  hipLaunchKernelGGL(k_zero_queue<0>, dim3(1), dim3(64), 0, st, p.queue);
  // 512 threads: 8 waves per CU with up to 256 VGPRs each
  int grid = run.grid > 0 ? run.grid : p.nrep;
  size_t lds = p.lds_resident ? (size_t)slab_doubles(p.max_n) * sizeof(double) : 0;
  if (p.lds_resident) {
    // the dynamic-LDS limit is raised once, outside any graph capture
    // (aw_resident_prepare)
    if (p.extras) hipLaunchKernelGGL((k_resident<T, true, 512, 3>), dim3(grid), dim3(512), lds, st, B, B.P, lg, mode);
    else if (p.table) hipLaunchKernelGGL((k_resident<T, true, 512, 1>), dim3(grid), dim3(512), lds, st, B, B.P, lg, mode);
    else hipLaunchKernelGGL((k_resident<T, true, 512, 0>), dim3(grid), dim3(512), lds, st, B, B.P, lg, mode);
  } else if (p.extras) {
    hipLaunchKernelGGL((k_resident<T, false, 512, 3>), dim3(grid), dim3(512), 0, st, B, B.P, lg, mode);
  } else if (p.table) {
    hipLaunchKernelGGL((k_resident<T, false, 512, 1>), dim3(grid), dim3(512), 0, st, B, B.P, lg, mode);
  } else {
    hipLaunchKernelGGL((k_resident<T, false, 512, 0>), dim3(grid), dim3(512), 0, st, B, B.P, lg, mode);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
