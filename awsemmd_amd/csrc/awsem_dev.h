// awsem_dev.h -- device-side batch description, replica views and the
// resident engine, shared by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "awsem_terms.h"
#include "awsem_launch.h"

// ------------------------------------------------------------------------
// energy epilogue: ered[N_EACC] (LDS) -> out[AW_N_ET]

// phase engine: blocks add their partial sums
__device__ __forceinline__ void energy_add(const double *ered, double *out)
{
  if (threadIdx.x < N_EACC) {
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
      for (int k = 0; k < N_EACC; ++k) v += ered[k];
    } else {
      for (int k = 0; k < N_EACC; ++k)
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
  V.ered = nullptr;
  V.P = (ParamPtr)B.P;
  V.fm_items = R.fm_items;
  V.fm_item_base = 0;
  V.coef = (const T *)R.coef;
  V.rf = R.rf;
  V.items = R.items;
  V.nl = nullptr;
  V.nlx = nullptr;
  V.nl_cap = 0;
  V.nl_state = nullptr;
  V.wq = nullptr;
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

template <class T, bool LDSM>
__device__ __forceinline__ void force_eval_block(const View<T> &V, unsigned long long *prof, bool first, bool ef)
{
  int t0 = threadIdx.x, ts = blockDim.x;
  unsigned long long t = prof ? wall_clock64() : 0;
  if (threadIdx.x < N_EACC) V.ered[threadIdx.x] = 0.0;
  phase_A(V, t0, ts);
  __syncthreads();
  prof_mark(prof, first, 0, t);
  if (V.nl) {
    bool build = nl_needs_build(V);
    if (build) phase_nl_build(V, t0, ts);
    __syncthreads();
    if (threadIdx.x == 0) nl_publish(V, build);
    __syncthreads();
  }
  prof_mark(prof, first, 11, t);
  // phase B in one part, or (clocked run) its three parts separately; one
  // call site keeps a single copy of the code
  int nparts = prof ? 3 : 1;
  for (int q = 0; q < nparts; ++q) {
    PairMarks pm{prof, first, t};
    phase_B<T, true, PairMarks, LDSM>(V, t0, ts, prof ? (1 << q) : 7, prof ? &pm : nullptr, ef);
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

template <class T, bool LDSM, int NT>
__global__ void __launch_bounds__(NT) k_resident(BatchDev B, const FFParams *__restrict__ Pr, LgParams lg, int mode)
{
  extern __shared__ double smem[];
  __shared__ int s_rep;
  __shared__ double ered[N_EACC];
  __shared__ __attribute__((aligned(8))) int nl_state[8];  // [2..5]: two 64-bit queue-counter words
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
      force_eval_block<T, LDSM>(V, B.prof, first, s == nst - 1);
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
  }
}

template <int D>
__global__ void k_zero_int(int *p)
{
  if (threadIdx.x == 0) *p = 0;
}

template <class T>
static void prepare_resident_t()
{
  int dev = 0, lim = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lim <= 0) lim = 65536;
  (void)hipFuncSetAttribute((const void *)k_resident<T, true, 512>, hipFuncAttributeMaxDynamicSharedMemorySize,
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
  hipLaunchKernelGGL(k_zero_int<0>, dim3(1), dim3(64), 0, st, p.queue);
  // 512 threads: 8 waves per CU with up to 256 VGPRs each
  int grid = run.grid > 0 ? run.grid : p.nrep;
  size_t lds = p.lds_resident ? (size_t)slab_doubles(p.max_n) * sizeof(double) : 0;
  if (p.lds_resident) {
    // the dynamic-LDS limit is raised once, outside any graph capture
    // (aw_resident_prepare)
    hipLaunchKernelGGL((k_resident<T, true, 512>), dim3(grid), dim3(512), lds, st, B, B.P, lg, mode);
  } else {
    hipLaunchKernelGGL((k_resident<T, false, 512>), dim3(grid), dim3(512), 0, st, B, B.P, lg, mode);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
