// awsem_server.cpp -- continuous batching of annealing jobs on one GPU.
//
// Jobs (a system, its start coordinates, a number of steps and a
// temperature ramp) queue on the host.  The scheduler works in rounds: it
// admits queued jobs in arrival order while job slots and the device-memory
// budget allow, uploads their device images, advances every active job in
// one resident-engine launch (jobs too large for LDS in a second launch on a
// second stream, overlapping the first) by a number of steps inversely
// proportional to its size, so that all workgroups finish together, then
// retires finished jobs -- final coordinates and per-term energies copied
// back, device memory freed -- which makes room for the next admissions.
// A job's coordinates, velocities, forces and Verlet list stay on the device
// between rounds, and its random stream is keyed by its own seed and step
// count, so its trajectory does not depend on which other jobs share the
// launches (up to the order of floating-point force accumulation).
#include "../../include/awsem.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "awsem_host.h"
#include "awsem_launch.h"

// defined in awsem_api.cpp
const AwSystem *aw_ctx_system(const awsem_ctx *c);

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b)
{
  return std::chrono::duration<double, std::milli>(b - a).count();
}

enum JobState { JOB_QUEUED = 0, JOB_RUNNING = 1, JOB_DONE = 2 };

struct Job {
  uint64_t id = 0;
  int n = 0;  // residues
  AwDevPlan plan;  // device layout and host image
  std::vector<double> x0, v0;
  int nsteps = 0, done = 0, round_steps = 0;
  double t_start = 300.0, t_stop = 300.0;
  uint64_t seed = 1;
  int state = JOB_QUEUED;  // or a negative AWSEM_E_* code
  std::string err;
  std::unique_ptr<AwDevSystem> dev;
  bool f_valid = false, fresh = true;
  Clock::time_point t_submit, t_admit, t_done;
  std::vector<double> x_out;
  double energy[AW_N_ET];
};

}  // namespace

struct awsem_server {
  FFParams P;
  awsem_server_opts o;
  int fp32 = 1;
  int lds_max = 0;
  size_t budget = 0, used = 0;
  hipStream_t st[2] = {nullptr, nullptr};
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  // steps of wide-engine jobs relative to the cost model, adapted each round
  // so that the two launches of a round take about the same time
  double wide_boost = 1.0;
  void *d_params = nullptr;
  ReplicaDesc *d_rep = nullptr;
  int *d_order = nullptr, *d_queue = nullptr;
  std::vector<ReplicaDesc> h_rep;
  std::vector<int> h_order;

  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job *> queued;
  std::vector<Job *> active;
  std::map<uint64_t, std::unique_ptr<Job>> jobs;
  uint64_t next_id = 1;
  std::thread worker;
  bool worker_on = false, stop_req = false;
  std::string err;
  // set when a scheduling round fails: the device state of every job in
  // flight is unknown, so the server refuses further work with this code
  int fail_code = 0;
  // test hook (AWSEM_SERVER_FAIL_ROUND=k): round k (1-based) fails before
  // it launches anything, as a failed launch would
  int fail_round = 0;
  // statistics
  double rounds = 0, jobs_done = 0, peak_active = 0, replica_steps = 0, device_ms = 0;

  ~awsem_server()
  {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop_req = true;
    }
    cv.notify_all();
    if (worker.joinable()) worker.join();
    for (auto &kv : jobs) kv.second->dev.reset();
    void *ptrs[] = {d_params, d_rep, d_order, d_queue};
    for (void *p : ptrs) if (p) (void)hipFree(p);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev2) (void)hipEventDestroy(ev2);
    if (ev3) (void)hipEventDestroy(ev3);
    for (hipStream_t s : st) if (s) (void)hipStreamDestroy(s);
  }

  AwStatus init(const FFParams &p, const awsem_server_opts &opts);
  AwStatus round(int *n_active, int *n_queued);
  void run_worker();
  // marks every job in flight or queued as failed with s (mu held)
  void fail_all(const AwStatus &s)
  {
    fail_code = s.code;
    err = s.msg;
    for (Job *j : active) { j->state = s.code; j->err = s.msg; j->t_done = Clock::now(); j->dev.reset(); }
    active.clear();
    used = 0;
    for (Job *j : queued) { j->state = s.code; j->err = s.msg; j->t_done = Clock::now(); }
    queued.clear();
  }
};

#define SV_HIPCHK(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return AwStatus::err(AWSEM_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

AwStatus awsem_server::init(const FFParams &p, const awsem_server_opts &opts)
{
  P = p;
  o = opts;
  fp32 = o.precision == AWSEM_PREC_FP32;
  lds_max = aw_lds_max_residues();
  aw_resident_prepare();
  if (o.hbm_budget_bytes > 0) {
    budget = (size_t)o.hbm_budget_bytes;
  } else {
    size_t fr = 0, tot = 0;
    SV_HIPCHK(hipMemGetInfo(&fr, &tot));
    budget = (size_t)(0.8 * (double)fr);
  }
  for (auto &s : st) SV_HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  SV_HIPCHK(hipEventCreate(&ev0));
  SV_HIPCHK(hipEventCreate(&ev1));
  SV_HIPCHK(hipEventCreate(&ev2));
  SV_HIPCHK(hipEventCreate(&ev3));
  AwStatus s = upload_raw(&d_params, &P, sizeof(FFParams));
  if (!s.ok()) return s;
  SV_HIPCHK(hipMalloc((void **)&d_rep, (size_t)o.max_jobs * sizeof(ReplicaDesc)));
  SV_HIPCHK(hipMalloc((void **)&d_order, (size_t)o.max_jobs * sizeof(int)));
  SV_HIPCHK(hipMalloc((void **)&d_queue, 256));
  h_rep.resize(o.max_jobs);
  h_order.resize(o.max_jobs);
  return AwStatus();
}

// one scheduling round: admit, launch, retire
AwStatus awsem_server::round(int *n_active, int *n_queued)
{
  std::vector<Job *> admit;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (fail_round > 0 && rounds + 1 >= fail_round)
      return AwStatus::err(AWSEM_E_DEVICE, "injected failure of scheduling round " + std::to_string(fail_round));
    while (!queued.empty() && (int)(active.size() + admit.size()) < o.max_jobs) {
      Job *j = queued.front();
      if (used + j->plan.bytes > budget) break;  // arrival order: wait for memory
      used += j->plan.bytes;
      queued.pop_front();
      admit.push_back(j);
    }
  }
  for (Job *j : admit) {
    std::unique_ptr<AwDevSystem> ds(new AwDevSystem());
    AwStatus s = aw_devsys_create(j->plan, *ds);
    j->plan.image = std::vector<unsigned char>();  // uploaded
    size_t bytes = (size_t)j->n * 9 * sizeof(double);
    if (s.ok() && (hipMemcpy(ds->d.x, j->x0.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
                   hipMemcpy(ds->d.v, j->v0.data(), bytes, hipMemcpyHostToDevice) != hipSuccess))
      s = AwStatus::err(AWSEM_E_DEVICE, "job upload");
    std::lock_guard<std::mutex> lk(mu);
    if (!s.ok()) {
      used -= j->plan.bytes;
      j->state = s.code;
      j->err = s.msg;
      j->t_done = Clock::now();
      cv.notify_all();
      continue;
    }
    j->dev = std::move(ds);
    j->state = JOB_RUNNING;
    j->t_admit = Clock::now();
    active.push_back(j);
  }
  std::vector<Job *> act;
  {
    std::lock_guard<std::mutex> lk(mu);
    act = active;
    peak_active = std::max(peak_active, (double)act.size());
  }
  if (!act.empty()) {
    // steps of each job this round: a job of median cost runs steps_per_round
    // steps, cheaper jobs proportionally more and dearer ones fewer, so that
    // every workgroup of the launch is busy for about the same time and small
    // jobs are not held back by the largest one (cost ~ residues, twice for
    // jobs working from global memory)
    std::vector<double> cost;
    for (Job *j : act) cost.push_back((double)j->n * (j->n <= lds_max ? 1.0 : 2.0));
    std::vector<double> sorted_cost = cost;
    std::nth_element(sorted_cost.begin(), sorted_cost.begin() + sorted_cost.size() / 2, sorted_cost.end());
    double c_med = sorted_cost[sorted_cost.size() / 2];
    for (size_t k = 0; k < act.size(); ++k) {
      Job *j = act[k];
      double want = o.steps_per_round * c_med / cost[k] * (j->n <= lds_max ? 1.0 : wide_boost);
      int steps = (int)std::lround(std::min(want, 8.0 * o.steps_per_round));
      j->round_steps = std::max(1, std::min(steps, j->nsteps - j->done));
    }
    // descriptors: LDS-resident jobs first, then the rest; longest first in each
    std::vector<Job *> grp[2];
    for (Job *j : act) grp[j->n <= lds_max ? 0 : 1].push_back(j);
    int base[2] = {0, (int)grp[0].size()};
    int max_n[2] = {0, 0};
    long long steps = 0;
    for (int g = 0; g < 2; ++g) {
      std::stable_sort(grp[g].begin(), grp[g].end(), [](Job *a, Job *b) { return a->n > b->n; });
      for (size_t k = 0; k < grp[g].size(); ++k) {
        Job *j = grp[g][k];
        ReplicaDesc R = j->dev->d;
        R.nsteps = j->round_steps;
        R.ramp0 = j->done;
        R.run_steps = j->nsteps;
        R.step0 = j->done;
        R.seed = j->seed;
        R.rng_id = 0;
        R.t_start = j->t_start;
        R.t_stop = j->t_stop;
        R.flags = (j->f_valid ? AW_RF_F_VALID : 0) | (j->fresh ? AW_RF_NL_RESET : 0);
        h_rep[base[g] + k] = R;
        h_order[base[g] + k] = (int)k;
        max_n[g] = std::max(max_n[g], j->n);
        steps += R.nsteps;
      }
    }
    int nact = (int)act.size();
    SV_HIPCHK(hipMemcpyAsync(d_rep, h_rep.data(), nact * sizeof(ReplicaDesc), hipMemcpyHostToDevice, st[0]));
    SV_HIPCHK(hipMemcpyAsync(d_order, h_order.data(), nact * sizeof(int), hipMemcpyHostToDevice, st[0]));
    SV_HIPCHK(hipStreamSynchronize(st[0]));
    SV_HIPCHK(hipEventRecord(ev0, st[0]));
    SV_HIPCHK(hipStreamWaitEvent(st[1], ev0, 0));
    for (int g = 0; g < 2; ++g) {
      if (grp[g].empty()) continue;
      AwBatchPtrs p;
      p.rep = d_rep + base[g];
      p.nrep = (int)grp[g].size();
      p.max_n = max_n[g];
      p.fp32 = fp32;
      p.lds_resident = g == 0 ? 1 : 0;
      p.table = P.fm_tb_flag;
      p.extras = P.xt_on;
      p.params = d_params;
      p.queue = d_queue + 32 * g;
      p.order = d_order + base[g];
      p.prof = nullptr;
      AwRun run;
      std::memset(&run, 0, sizeof run);
      run.mode = 1;
      run.dt = o.dt;
      run.damp = o.damp;
      // jobs too large for LDS: the wide engine spreads each over many CUs
      int max_steps = 0;
      for (Job *j : grp[g]) max_steps = std::max(max_steps, j->round_steps);
      int lrc = g == 0 ? aw_launch_resident(p, run, st[g]) : aw_launch_wide(p, run, max_steps, st[g]);
      if (lrc) return AwStatus::err(AWSEM_E_DEVICE, "server launch failed");
    }
    SV_HIPCHK(hipEventRecord(ev3, st[1]));
    SV_HIPCHK(hipEventRecord(ev2, st[0]));
    SV_HIPCHK(hipStreamWaitEvent(st[1], ev2, 0));
    SV_HIPCHK(hipEventRecord(ev1, st[1]));
    SV_HIPCHK(hipStreamSynchronize(st[1]));
    float ms = 0.f, ms_res = 0.f, ms_wide = 0.f;
    (void)hipEventElapsedTime(&ms, ev0, ev1);
    (void)hipEventElapsedTime(&ms_res, ev0, ev2);
    (void)hipEventElapsedTime(&ms_wide, ev0, ev3);
    if (!grp[1].empty()) {
      double r = grp[0].empty() ? 2.0 : (double)ms_res / std::max(1e-3, (double)ms_wide);
      wide_boost = std::min(16.0, std::max(0.25, wide_boost * std::min(2.0, std::max(0.5, r))));
    }
    // retire
    std::vector<Job *> finished;
    for (Job *j : act) {
      j->done += j->round_steps;
      j->f_valid = true;
      j->fresh = false;
      if (j->done >= j->nsteps) finished.push_back(j);
    }
    for (Job *j : finished) {
      size_t bytes = (size_t)j->n * 9 * sizeof(double);
      j->x_out.resize((size_t)j->n * 9);
      AwStatus s;
      if (hipMemcpy(j->x_out.data(), j->dev->d.x, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(j->energy, j->dev->d.energy, sizeof j->energy, hipMemcpyDeviceToHost) != hipSuccess)
        s = AwStatus::err(AWSEM_E_DEVICE, "job result copy");
      // table mode: a distance outside the table's range in any round of
      // the job (the error word is never cleared during a job)
      if (s.ok() && j->dev->d.tb_items > 0) {
        int w = 0;
        if (hipMemcpy(&w, j->dev->d.fault, sizeof w, hipMemcpyDeviceToHost) != hipSuccess)
          s = AwStatus::err(AWSEM_E_DEVICE, "job error word");
        else if (w & AW_FAULT_TB_RANGE)
          s = AwStatus::err(AWSEM_E_RANGE, "Table Fragment Memory: r is out of computed range.");
      }
      j->dev.reset();
      std::lock_guard<std::mutex> lk(mu);
      used -= j->plan.bytes;
      j->state = s.ok() ? JOB_DONE : s.code;
      j->err = s.msg;
      j->t_done = Clock::now();
      active.erase(std::find(active.begin(), active.end(), j));
      jobs_done += 1;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      rounds += 1;
      replica_steps += (double)steps;
      device_ms += ms;
    }
    cv.notify_all();
  }
  std::lock_guard<std::mutex> lk(mu);
  if (n_active) *n_active = (int)active.size();
  if (n_queued) *n_queued = (int)queued.size();
  return AwStatus();
}

void awsem_server::run_worker()
{
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return stop_req || !queued.empty() || !active.empty(); });
      if (stop_req) return;
    }
    AwStatus s = round(nullptr, nullptr);
    if (!s.ok()) {
      std::lock_guard<std::mutex> lk(mu);
      fail_all(s);
      // the thread ends here: the server no longer counts as started, and
      // every later call reports the failure (worker is joined by stop,
      // start or the destructor)
      worker_on = false;
      cv.notify_all();
      return;
    }
  }
}

static int sfail(awsem_server *s, const AwStatus &st)
{
  if (s) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->err = st.msg;
  }
  return st.code;
}

extern "C" {

int awsem_server_create(const awsem_ctx *forcefield, const awsem_server_opts *opts, awsem_server **out)
{
  if (!out || !forcefield || !opts) return AWSEM_E_ARG;
  *out = nullptr;
  const AwSystem *S = aw_ctx_system(forcefield);
  if (!S || S->n < 0) return AWSEM_E_ARG;
  if (opts->max_jobs <= 0 || opts->steps_per_round <= 0 || !(opts->dt > 0) || !(opts->damp > 0) ||
      (opts->precision != AWSEM_PREC_FP32 && opts->precision != AWSEM_PREC_FP64))
    return AWSEM_E_ARG;
  std::unique_ptr<awsem_server> s(new awsem_server());
  if (const char *e = std::getenv("AWSEM_SERVER_FAIL_ROUND")) s->fail_round = std::atoi(e);
  AwStatus st = s->init(S->P, *opts);
  *out = s.release();
  if (!st.ok()) { (*out)->err = st.msg; return st.code; }
  return AWSEM_OK;
}

int awsem_server_submit(awsem_server *s, const awsem_ctx *system, const double *xres, const double *vres, int nsteps,
                        double t_start, double t_stop, uint64_t seed, uint64_t *job_id)
{
  if (!s || !system || !xres || nsteps <= 0 || !job_id) return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(system);
  if (!S || S->n <= 0) return AWSEM_E_ARG;
  if (S->n >= 32768) return sfail(s, AwStatus::err(AWSEM_E_ARG, "served systems are limited to 32767 residues each"));
  std::unique_ptr<Job> j(new Job());
  j->n = S->n;
  size_t n9 = (size_t)S->n * 9;
  for (size_t k = 0; k < n9; ++k)
    if (!std::isfinite(xres[k]) || (vres && !std::isfinite(vres[k])))
      return sfail(s, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates or velocities"));
  j->x0.assign(xres, xres + n9);
  j->v0.assign(n9, 0.0);
  if (vres) j->v0.assign(vres, vres + n9);
  j->nsteps = nsteps;
  j->t_start = t_start;
  j->t_stop = t_stop;
  j->seed = seed;
  // the server's force field applies to every job; the image is built here,
  // on the submitting thread, so that admission is an allocation and a copy
  AwStatus st = aw_devsys_plan(*S, &s->P, s->fp32, S->n > s->lds_max, true, j->plan);
  if (!st.ok()) return sfail(s, st);
  if (j->plan.bytes > s->budget) return sfail(s, AwStatus::err(AWSEM_E_NOMEM, "job exceeds the device-memory budget"));
  j->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fail_code) return s->fail_code;  // s->err holds the failure
    j->id = s->next_id++;
    *job_id = j->id;
    s->queued.push_back(j.get());
    s->jobs[j->id] = std::move(j);
  }
  s->cv.notify_all();
  return AWSEM_OK;
}

int awsem_server_pump(awsem_server *s, int *n_active, int *n_queued)
{
  if (!s) return AWSEM_E_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fail_code) return s->fail_code;
    if (s->worker_on) return AWSEM_E_STATE;
  }
  AwStatus st = s->round(n_active, n_queued);
  if (!st.ok()) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->fail_all(st);
    s->cv.notify_all();
    return st.code;
  }
  return AWSEM_OK;
}

int awsem_server_start(awsem_server *s)
{
  if (!s) return AWSEM_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->fail_code) return s->fail_code;
  if (s->worker_on) return AWSEM_E_STATE;
  if (s->worker.joinable()) {
    // a worker that ended on its own (a failed round) is joined first
    lk.unlock();
    s->worker.join();
    lk.lock();
  }
  s->stop_req = false;
  s->worker_on = true;
  s->worker = std::thread([s] { s->run_worker(); });
  return AWSEM_OK;
}

int awsem_server_stop(awsem_server *s)
{
  if (!s) return AWSEM_E_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop_req = true;
  }
  s->cv.notify_all();
  // also joins a worker that ended on its own after a failed round
  if (s->worker.joinable()) s->worker.join();
  std::lock_guard<std::mutex> lk(s->mu);
  s->worker_on = false;
  s->stop_req = false;
  return AWSEM_OK;
}

int awsem_server_poll(awsem_server *s, uint64_t job_id, int *state)
{
  if (!s || !state) return AWSEM_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return AWSEM_E_ARG;
  *state = it->second->state;
  return AWSEM_OK;
}

int awsem_server_wait(awsem_server *s, uint64_t job_id, double timeout_s)
{
  if (!s) return AWSEM_E_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return AWSEM_E_ARG;
  Job *j = it->second.get();
  // a failed server fails every job it held; a job still queued or running
  // there could otherwise never finish
  auto done = [&] { return j->state == JOB_DONE || j->state < 0 || s->fail_code != 0; };
  if (timeout_s > 0) {
    if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return AWSEM_E_STATE;
  } else {
    s->cv.wait(lk, done);
  }
  if (j->state < 0) return j->state;
  return j->state == JOB_DONE ? AWSEM_OK : s->fail_code;
}

int awsem_server_result(awsem_server *s, uint64_t job_id, double *xres, double *energy, double *times_ms)
{
  if (!s) return AWSEM_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return AWSEM_E_ARG;
  Job *j = it->second.get();
  if (j->state < 0) { s->err = j->err; return j->state; }
  if (j->state != JOB_DONE) return AWSEM_E_STATE;
  if (xres) std::memcpy(xres, j->x_out.data(), j->x_out.size() * sizeof(double));
  if (energy) std::memcpy(energy, j->energy, sizeof j->energy);
  if (times_ms) {
    times_ms[0] = ms_since(j->t_submit, j->t_admit);
    times_ms[1] = ms_since(j->t_admit, j->t_done);
    times_ms[2] = ms_since(j->t_submit, j->t_done);
  }
  return AWSEM_OK;
}

int awsem_server_release(awsem_server *s, uint64_t job_id)
{
  if (!s) return AWSEM_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return AWSEM_E_ARG;
  if (it->second->state == JOB_QUEUED || it->second->state == JOB_RUNNING) return AWSEM_E_STATE;
  s->jobs.erase(it);
  return AWSEM_OK;
}

int awsem_server_stats(awsem_server *s, double *stats)
{
  if (!s || !stats) return AWSEM_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  stats[0] = s->rounds;
  stats[1] = s->jobs_done;
  stats[2] = (double)s->used;
  stats[3] = s->peak_active;
  stats[4] = s->replica_steps;
  stats[5] = s->device_ms;
  stats[6] = (double)s->budget;
  stats[7] = (double)s->active.size();
  return AWSEM_OK;
}

int awsem_server_last_error(awsem_server *s, char *buf, size_t n)
{
  if (!s || !buf || n == 0) return AWSEM_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  std::snprintf(buf, n, "%s", s->err.c_str());
  return AWSEM_OK;
}

int awsem_server_destroy(awsem_server *s)
{
  delete s;
  return AWSEM_OK;
}

}  // extern "C"
