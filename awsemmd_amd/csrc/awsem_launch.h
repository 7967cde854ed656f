// awsem_launch.h -- internal C++ interface between the host API and the
// gfx950 kernels (not part of the public C ABI).
#pragma once
#include <hip/hip_runtime.h>

struct AwBatchPtrs {
  const void *rep;      // ReplicaDesc[nrep]
  int nrep;
  int max_n;            // largest replica (residues)
  int fp32;             // pair arithmetic in FP32 (else FP64)
  int lds_resident;     // resident engine keeps replicas in LDS
  const void *params;   // FFParams
  const int *rinfo;
  const double *aps;
  const int *ch_start, *ch_len;
  double *x, *v, *f, *energy, *scratch;
  const void *coef, *rf, *items;
  int *queue;
  const int *order;
  unsigned long long *prof;  // optional per-phase clock accumulators [32]
  int *nl;                   // Verlet lists of the resident engine
  double *nlx;
};

struct AwRun {
  int mode;             // 0 force evaluation, 1 dynamics
  int nsteps;
  double dt, t_start, t_stop, damp;
  unsigned long long seed;
  long long step0;
  int threads, grid;    // 0 = defaults
  int nl_reset;         // state replaced since the last launch
  int f_valid;          // device forces belong to the current coordinates
};

int aw_lds_max_residues();
long long aw_slab_doubles(int n);
int aw_launch_resident(const AwBatchPtrs &p, const AwRun &run, hipStream_t st);
int aw_launch_phases(const AwBatchPtrs &p, const AwRun &run, hipStream_t st);
int aw_launch_list(const AwBatchPtrs &p, int r, int kind, int arg, long long *out, int cap, int *count, hipStream_t st);
