// awsem_kernels.hip -- gfx950 kernels of the FixBackbone force path.
//
// One force evaluation is a fixed sequence of phases over one replica
// (a system of one or more chains):
//   A  virtual atoms N, H, C' (fb.cpp:7113-7138), per-residue extents,
//      zero accumulators
//   B  fragment memory (fb.cpp:4079-4177), chain/chi/rama (2062-2546),
//      harmonic bonds, and one sweep over residue pairs for the density
//      rho (DL1, 7729-7786), DSSP (2902-3277), P_AP (3279-3383) and the
//      excluded-volume pair style (pev.cpp:50-138)
//   C  sigma_h and burial (SL, 7795-7841)
//   D  helix (i,i+4) pairs (7843-7868, 3706-3785) and water xi (DL2, 7879-7937)
//   E  xi *= dsigma/drho (7939-7950)
//   F  water / burial / helix pair forces (DL3, 7996-8057)
// Coeff-toggled terms (awsem_extra.h) run in their own pass after A.
// Two engines run the same phase bodies:
//   * resident: one workgroup owns one replica, its coordinates, forces and
//     per-residue scalars live in LDS (or in a global scratch slab for
//     replicas too large for LDS), phases are separated by workgroup
//     barriers and nsteps of FP64 Langevin dynamics run inside one launch;
//   * phase: every phase is its own grid-wide launch over one replica's
//     items with state in global memory (the plain single-chain drop-in).
// Pair and switching arithmetic runs in T (float or double); coordinates,
// force accumulation, energies, list predicates and the integrator are FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "awsem_dev.h"


// ------------------------------------------------------------------------
// phase engine: each phase is a grid-wide launch over one replica

template <class T>
__device__ __forceinline__ View<T> phase_view(const BatchDev &B, int r)
{
  const ReplicaDesc &R = B.rep[r];
  return make_view<T>(B, R, R.scratch);
}

template <class T>
__global__ void __launch_bounds__(256) k_ph_load(BatchDev B, int r)
{
  const ReplicaDesc &R = B.rep[r];
  View<T> V = phase_view<T>(B, r);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 9 * R.n; k += gridDim.x * blockDim.x) V.x()[k] = R.x[k];
}

template <class T>
__global__ void __launch_bounds__(256) k_ph_A(BatchDev B, int r)
{
  View<T> V = phase_view<T>(B, r);
  phase_A(V, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
  if (blockIdx.x == 0 && threadIdx.x < AW_N_ET) B.rep[r].energy[threadIdx.x] = 0.0;
}

#define PH_KERNEL(NAME, CALL)                                                          \
  template <class T>                                                                   \
  __global__ void __launch_bounds__(256) NAME(BatchDev B, int r)                       \
  {                                                                                    \
    __shared__ double ered[N_ERED];                                                    \
    View<T> V = phase_view<T>(B, r);                                                   \
    V.ered = ered;                                                                     \
    if (threadIdx.x < N_ERED) ered[threadIdx.x] = 0.0;                                 \
    __syncthreads();                                                                   \
    CALL;                                                                              \
    __syncthreads();                                                                   \
    energy_add(ered, B.rep[r].energy);                                                 \
  }

PH_KERNEL(k_ph_X, xt_pass(V, B.rep[r], blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, true))
PH_KERNEL(k_ph_B, (phase_B<T, false>(V, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, 7)))
PH_KERNEL(k_ph_C, phase_C(V, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x))
PH_KERNEL(k_ph_D, phase_D(V, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x))
PH_KERNEL(k_ph_F, phase_F(V, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x))

template <class T>
__global__ void __launch_bounds__(256) k_ph_store(BatchDev B, int r, int with_x)
{
  const ReplicaDesc &R = B.rep[r];
  View<T> V = phase_view<T>(B, r);
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 9 * R.n; k += gridDim.x * blockDim.x) {
    R.f[k] = V.f()[k];
    if (with_x) R.x[k] = V.x()[k];
  }
}

template <class T>
__global__ void __launch_bounds__(256) k_ph_int1(BatchDev B, int r, LgParams lg)
{
  View<T> V = phase_view<T>(B, r);
  integrate1(V, (AW_G double *)B.rep[r].v, lg, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

template <class T>
__global__ void __launch_bounds__(256) k_ph_int2(BatchDev B, int r, LgParams lg, int s)
{
  const ReplicaDesc R = B.rep[r];
  View<T> V = phase_view<T>(B, r);
  integrate2(V, (AW_G double *)R.v, lg, R, s, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// total energy slot of one replica
__global__ void k_energy_total(BatchDev B, int r)
{
  if (threadIdx.x != 0) return;
  double *e = B.rep[r].energy;
  double s = 0.0;
  for (int k = 1; k < AW_N_ET; ++k) s += e[k];
  e[0] = s;
}

// ------------------------------------------------------------------------
// integer lists for parity tests (same predicates as the force passes)

template <class T>
__global__ void __launch_bounds__(256) k_list(BatchDev B, int r, int kind, int arg, long long *out, int cap, int *count)
{
  View<T> V = phase_view<T>(B, r);
  ParamPtr P = V.P;
  int n = V.n;
  int gt = blockIdx.x * blockDim.x + threadIdx.x, gs = gridDim.x * blockDim.x;
  if (kind == 4) {
    for (int i = gt; i < n; i += gs) {
      int j = i + P->helix_i_diff;
      if (!P->helix_flag || j >= n) continue;
      if (rchain(V.rinfo[i]) != rchain(V.rinfo[j])) continue;
      double r2 = rsq_exact(XA(V, K_O, i), XV(V, K_N, j));
      if (r2 < P->helix_cutoff_sq && !(sqrt(r2) > P->helix_cutoff)) {
        int k = atomicAdd(count, 1);
        if (k < cap) out[k] = (long long)i * n + j;
      }
    }
    return;
  }
  int np = pair_count(n);
  for (int k = gt; k < np; k += gs) {
    int i, j;
    pair_decode(k, n, i, j);
    int ri = V.rinfo[i], rj = V.rinfo[j];
    bool same = rchain(ri) == rchain(rj);
    int dres = j - i;
    long long code = -1, code2 = -1;
    if (kind == 0 || kind == 1 || kind == 6) {
      double rsq = rsq_exact(site(V, i, ri), site(V, j, rj));
      if (kind == 0 && P->water_flag && (!same || dres > 1) && arg < P->n_wells && rsq > P->w_rmin_th_sq[arg] && rsq < P->w_rmax_th_sq[arg])
        code = (long long)i * n + j;
      if (kind == 1 && P->helix_flag && (!same || dres > 1) && rsq > P->h_rmin_th_sq && rsq < P->h_rmax_th_sq)
        code = (long long)i * n + j;
      if (kind == 6 && P->water_flag && (!same || dres >= P->contact_cutoff) && arg < P->n_wells && P->well_flag[arg]) {
        int it = rtype(ri), jt = rtype(rj);
        const auto wg = ((ri | rj) & RI_PHOS) ? P->phosph_gamma : P->water_gamma;
        double g0 = wg[arg][it][jt][0], g1 = wg[arg][it][jt][1];
        if (((double)(!fabs(g0 - g1)) < AW_DELTA) && rsq > P->w_rmin_th_sq[arg] && rsq < P->w_rmax_th_sq[arg])
          code = (long long)i * n + j;
      }
    } else if (kind == 2) {
      if (P->dssp_flag && (!same || dres > 2)) {
        if (!(rj & RI_PRO) && !(ri & RI_LAST) && !(rj & RI_FIRST)) {
          double r2 = rsq_exact(XA(V, K_O, i), XV(V, K_N, j));
          if (r2 < P->dssp_cut_sq && !(sqrt(r2) > P->dssp_cut)) code = (long long)i * n + j;
        }
        if (!(ri & RI_PRO) && !(rj & RI_LAST) && !(ri & RI_FIRST)) {
          double r2 = rsq_exact(XA(V, K_O, j), XV(V, K_N, i));
          if (r2 < P->dssp_cut_sq && !(sqrt(r2) > P->dssp_cut)) code2 = (long long)j * n + i;
        }
      }
    } else if (kind == 3) {
      if (P->pap_flag && rsq_exact(XA(V, K_CA, i), XA(V, K_CA, j)) < P->pap_cutoff_sq) code = (long long)i * n + j;
    } else if (kind == 5) {
      bool nb = same && (j == i + 1), shortc = same && (j - i < 5);
      int na = 3 * n;
      for (int ka = 0; ka < 3; ++ka)
        for (int kb = 0; kb < 3; ++kb) {
          if (nb && kb == K_CA && (ka == K_CA || ka == K_O)) continue;
          int ta = atype(ka, ri), tb = atype(kb, rj);
          if (P->ev_lambda[ta][tb] == 0.0) continue;
          double rc = shortc ? P->ev_cut_short[ta][tb] : P->ev_cut_long[ta][tb];
          double rr = sqrt(rsq_exact(XA(V, ka, i), XA(V, kb, j)));
          if (rr < rc) {
            long long a = 3LL * i + ka, b = 3LL * j + kb;
            long long cc = a < b ? a * na + b : b * na + a;
            int q = atomicAdd(count, 1);
            if (q < cap) out[q] = cc;
          }
        }
    }
    if (code >= 0) { int q = atomicAdd(count, 1); if (q < cap) out[q] = code; }
    if (code2 >= 0) { int q = atomicAdd(count, 1); if (q < cap) out[q] = code2; }
  }
  if (kind == 5 && gt < n) {
    // same-residue CB-O
    int i = gt;
    int ri = V.rinfo[i];
    int ta = atype(K_CB, ri);
    if (P->ev_lambda[ta][3] != 0.0) {
      double rr = sqrt(rsq_exact(XA(V, K_CB, i), XA(V, K_O, i)));
      if (rr < P->ev_cut_short[ta][3]) {
        int na = 3 * n;
        int q = atomicAdd(count, 1);
        if (q < cap) out[q] = (3LL * i + 1) * na + 3LL * i + 2;
      }
    }
  }
}

// ------------------------------------------------------------------------
// host-side launchers (internal C++ interface used by awsem_api.cpp)

#include "awsem_launch.h"

static int g_lds_max = 0, g_lds_lim = 0;

int aw_lds_limit_bytes()
{
  if (!g_lds_lim) {
    int dev = 0, lim = 0;
    hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lim <= 0) lim = 65536;
    g_lds_lim = lim;
  }
  return g_lds_lim;
}

int aw_lds_max_residues()
{
  if (!g_lds_max) {
    int lim = aw_lds_limit_bytes();
    long long avail = (lim - AW_LDS_STATIC_RESERVE) / 8;  // dynamic LDS, doubles
    int n = 1;
    while (slab_doubles(n + 1) <= avail) ++n;
    g_lds_max = n;
  }
  return g_lds_max;
}

long long aw_slab_doubles(int n) { return slab_doubles(n); }

int aw_launch_resident_f32(const AwBatchPtrs &p, const AwRun &run, hipStream_t st);
int aw_launch_resident_f64(const AwBatchPtrs &p, const AwRun &run, hipStream_t st);

void aw_resident_prepare_f32();
void aw_resident_prepare_f64();
void aw_wide_prepare_f32();
void aw_wide_prepare_f64();

void aw_resident_prepare()
{
  static bool done = false;
  if (done) return;
  aw_resident_prepare_f32();
  aw_resident_prepare_f64();
  aw_wide_prepare_f32();
  aw_wide_prepare_f64();
  done = true;
}

int aw_launch_resident(const AwBatchPtrs &p, const AwRun &run, hipStream_t st)
{
  return p.fp32 ? aw_launch_resident_f32(p, run, st) : aw_launch_resident_f64(p, run, st);
}

template <class T>
static void force_phases(const BatchDev &B, int r, int grid, hipStream_t st, bool xt)
{
  hipLaunchKernelGGL(k_ph_A<T>, dim3(grid), dim3(256), 0, st, B, r);
  if (xt) hipLaunchKernelGGL(k_ph_X<T>, dim3(grid), dim3(256), 0, st, B, r);
  hipLaunchKernelGGL(k_ph_B<T>, dim3(grid), dim3(256), 0, st, B, r);
  hipLaunchKernelGGL(k_ph_C<T>, dim3(grid), dim3(256), 0, st, B, r);
  hipLaunchKernelGGL(k_ph_D<T>, dim3(grid), dim3(256), 0, st, B, r);
  hipLaunchKernelGGL(k_ph_F<T>, dim3(grid), dim3(256), 0, st, B, r);
  hipLaunchKernelGGL(k_energy_total, dim3(1), dim3(64), 0, st, B, r);
}

// the phase engine: replicas one after another, each phase a grid-wide
// launch; dynamics follow the replicas' descriptors (schedule, seed)
template <class T>
static int launch_phases_t(const AwBatchPtrs &p, const AwRun &run, hipStream_t st)
{
  BatchDev B = to_dev(p);
  LgParams lg = lg_make(run.dt, run.damp);
  int grid = run.grid > 0 ? run.grid : 1024;
  for (int r = 0; r < p.nrep; ++r) {
    hipLaunchKernelGGL(k_ph_load<T>, dim3(grid), dim3(256), 0, st, B, r);
    force_phases<T>(B, r, grid, st, p.extras != 0);
    if (run.mode == 1) {
      for (int s = 0; s < run.nsteps; ++s) {
        hipLaunchKernelGGL(k_ph_int1<T>, dim3(grid), dim3(256), 0, st, B, r, lg);
        force_phases<T>(B, r, grid, st, p.extras != 0);
        hipLaunchKernelGGL(k_ph_int2<T>, dim3(grid), dim3(256), 0, st, B, r, lg, s);
      }
    }
    hipLaunchKernelGGL(k_ph_store<T>, dim3(grid), dim3(256), 0, st, B, r, run.mode == 1 ? 1 : 0);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int aw_launch_wide_f32(const AwBatchPtrs &p, const AwRun &run, int max_steps, hipStream_t st);
int aw_launch_wide_f64(const AwBatchPtrs &p, const AwRun &run, int max_steps, hipStream_t st);

int aw_launch_wide(const AwBatchPtrs &p, const AwRun &run, int max_steps, hipStream_t st)
{
  return p.fp32 ? aw_launch_wide_f32(p, run, max_steps, st) : aw_launch_wide_f64(p, run, max_steps, st);
}

int aw_launch_phases(const AwBatchPtrs &p, const AwRun &run, hipStream_t st)
{
  return p.fp32 ? launch_phases_t<float>(p, run, st) : launch_phases_t<double>(p, run, st);
}

int aw_launch_list(const AwBatchPtrs &p, int r, int kind, int arg, long long *out, int cap, int *count, hipStream_t st)
{
  BatchDev B = to_dev(p);
  hipMemsetAsync(count, 0, sizeof(int), st);
  // virtual atoms first (phase A of this replica)
  if (p.fp32) {
    hipLaunchKernelGGL(k_ph_load<float>, dim3(256), dim3(256), 0, st, B, r);
    hipLaunchKernelGGL(k_ph_A<float>, dim3(256), dim3(256), 0, st, B, r);
    hipLaunchKernelGGL(k_list<float>, dim3(512), dim3(256), 0, st, B, r, kind, arg, out, cap, count);
  } else {
    hipLaunchKernelGGL(k_ph_load<double>, dim3(256), dim3(256), 0, st, B, r);
    hipLaunchKernelGGL(k_ph_A<double>, dim3(256), dim3(256), 0, st, B, r);
    hipLaunchKernelGGL(k_list<double>, dim3(512), dim3(256), 0, st, B, r, kind, arg, out, cap, count);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------------------
// order parameters (compute qwolynes, compute qonuchic): one workgroup per
// trace sums its pair items in a fixed order (lane-strided, then a tree in
// LDS), so a value does not depend on what else is launched

__global__ void __launch_bounds__(1024) k_q(const QItem *__restrict__ items, int nitems, int gauss,
                                           const double *__restrict__ xbase, const long long *__restrict__ xoff,
                                           double norm, double *__restrict__ out)
{
  __shared__ double part[1024];
  const double *x = xbase + xoff[blockIdx.x];
  double s = 0.0;
  for (int k = threadIdx.x; k < nitems; k += blockDim.x) {
    QItem it = items[k];
    const double *a = x + 3 * it.i, *b = x + 3 * it.j;
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    double r2 = dx * dx + dy * dy + dz * dz;
    if (gauss) {
      double d = sqrt(r2) - it.a;
      s += exp(-d * d * it.b);
    } else if (r2 < it.a) {
      s += 1.0;
    }
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = part[0] * norm;
}

int aw_launch_q(const void *items, int nitems, int gauss, const double *xbase, const long long *xoff, int count,
                double norm, double *out, hipStream_t st)
{
  if (count <= 0) return 0;
  hipLaunchKernelGGL(k_q, dim3(count), dim3(1024), 0, st, (const QItem *)items, nitems, gauss, xbase, xoff, norm, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
