// awsem_launch.h -- internal C++ interface between the host API and the
// gfx950 kernels (not part of the public C ABI).
#pragma once
#include <hip/hip_runtime.h>

struct AwBatchPtrs {
  const void *rep = nullptr;      // ReplicaDesc[nrep] on the device
  int nrep = 0;
  int max_n = 0;                  // largest replica (residues)
  int fp32 = 0;                   // pair arithmetic in FP32 (else FP64)
  int lds_resident = 0;           // resident engine keeps replicas in LDS
  int table = 0;                  // fragment-memory table mode (the resident engine's
                                  // instance with the table pass; the others have none)
  int extras = 0;                 // coeff-toggled terms on (the engines' extra pass)
  const void *params = nullptr;   // FFParams on the device
  int *queue = nullptr;           // device scratch: [0] work-queue counter, [16] replicas finished,
                                  // [32] sharing error flag, [48] chunks evaluated elsewhere (the
                                  // last three with nshare > 0 only)
  const int *order = nullptr;     // [nrep] replica order for the resident engine
  unsigned long long *prof = nullptr;  // optional per-phase clock accumulators [32]
  // fragment-memory sharing (resident engine, LDS variant): the replicas
  // whose chunks other workgroups may take
  int nshare = 0;
  const int *share_rep = nullptr;             // [nshare] replica index
  const int *share_n = nullptr;               // [nshare] chunks
  unsigned long long *share_word = nullptr;   // [nshare] claim words (ReplicaDesc::sh_word)
  int share_spread = 1;                       // helpers start their scan at blockIdx mod this (power of 2, <= 64)
};

struct AwRun {
  int mode;             // 0 force evaluation, 1 dynamics
  int nsteps;           // phase engine: steps of this launch
  double dt, damp;
  int grid;             // 0 = default
};

int aw_lds_max_residues();
// bytes of LDS one workgroup may use (the device's limit)
int aw_lds_limit_bytes();
long long aw_slab_doubles(int n);
// raises the resident and wide kernels' dynamic-LDS limits (once per
// process, before any launch or graph capture)
void aw_resident_prepare();
int aw_launch_resident(const AwBatchPtrs &p, const AwRun &run, hipStream_t st);
int aw_launch_phases(const AwBatchPtrs &p, const AwRun &run, hipStream_t st);
// the wide engine: every phase one launch over all replicas (working slabs,
// lists and queues in global memory); max_steps = largest replica nsteps
int aw_launch_wide(const AwBatchPtrs &p, const AwRun &run, int max_steps, hipStream_t st);
int aw_launch_list(const AwBatchPtrs &p, int r, int kind, int arg, long long *out, int cap, int *count, hipStream_t st);

// the residue-layout evaluation over a caller-given device region
// (awsem_compute_res_noalloc): doubles the region needs for n residues; the
// checks of a call (arguments, size against cap_doubles, device image built,
// finite coordinates; 0 or an AWSEM_E_* code with the context's message
// set); the evaluation itself, fres += forces
struct awsem_ctx;
size_t aw_noalloc_doubles(int n);
int aw_noalloc_check(awsem_ctx *c, const double *xres, const double *fres, size_t cap_doubles);
int aw_compute_res_scratch(awsem_ctx *c, const double *xres, double *fres, double *energy, double *dscr);

// order parameters: Q of `count` CA traces, trace r at xbase + xoff[r]
// ([n][3] doubles); items: QItem[nitems]; gauss: 1 for the Gaussian kinds,
// 0 for the contact counts; out[r] = norm * sum (device pointers)
int aw_launch_q(const void *items, int nitems, int gauss, const double *xbase, const long long *xoff, int count,
                double norm, double *out, hipStream_t st);
// the order-parameter object (awsem_q.cpp): Q of `count` CA traces at
// d_xbase + xoff[r] (host offsets, doubles) into values (host), on stream st
struct awsem_q;
int aw_q_eval(awsem_q *q, const double *d_xbase, const long long *xoff, int count, double *values, hipStream_t st);
int aw_q_residues(const awsem_q *q);
