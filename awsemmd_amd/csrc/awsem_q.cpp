// awsem_q.cpp -- order parameters of a CA trace on the device: compute
// qwolynes (src/compute_q_wolynes.cpp) and compute qonuchic
// (src/compute_q_onuchic.cpp) as pair lists evaluated by one workgroup per
// trace, for one trace or for the replicas of a batch in place.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/awsem.h"
#include "awsem_host.h"
#include "awsem_launch.h"

struct awsem_q {
  int kind = 0, n = 0;
  double norm = 0.0;
  std::vector<QItem> items;
  std::string err;
  void *d_items = nullptr;
  double *d_x = nullptr, *d_out = nullptr;
  long long *d_off = nullptr;
  int cap = 0;  // traces d_off / d_out hold
  ~awsem_q()
  {
    void *p[] = {d_items, d_x, d_out, d_off};
    for (void *q : p)
      if (q) (void)hipFree(q);
  }
};

static int qfail(awsem_q *q, int code, const std::string &m)
{
  q->err = m;
  return code;
}

// the items on the device and room for `count` traces
static int q_ready(awsem_q *q, int count)
{
  if (!q->d_items) {
    size_t b = std::max<size_t>(q->items.size(), 1) * sizeof(QItem);
    if (hipMalloc(&q->d_items, b) != hipSuccess ||
        (!q->items.empty() && hipMemcpy(q->d_items, q->items.data(), q->items.size() * sizeof(QItem),
                                        hipMemcpyHostToDevice) != hipSuccess))
      return qfail(q, AWSEM_E_DEVICE, "order parameter: device upload failed");
    if (hipMalloc((void **)&q->d_x, (size_t)q->n * 3 * sizeof(double)) != hipSuccess)
      return qfail(q, AWSEM_E_DEVICE, "order parameter: device allocation failed");
  }
  if (count > q->cap) {
    if (q->d_out) (void)hipFree(q->d_out);
    if (q->d_off) (void)hipFree(q->d_off);
    q->d_out = nullptr;
    q->d_off = nullptr;
    if (hipMalloc((void **)&q->d_out, (size_t)count * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&q->d_off, (size_t)count * sizeof(long long)) != hipSuccess)
      return qfail(q, AWSEM_E_DEVICE, "order parameter: device allocation failed");
    q->cap = count;
  }
  return AWSEM_OK;
}

int aw_q_eval(awsem_q *q, const double *d_xbase, const long long *xoff, int count, double *values, hipStream_t st)
{
  int rc = q_ready(q, count);
  if (rc) return rc;
  if (hipMemcpyAsync(q->d_off, xoff, (size_t)count * sizeof(long long), hipMemcpyHostToDevice, st) != hipSuccess ||
      aw_launch_q(q->d_items, (int)q->items.size(), q->kind == AWSEM_Q_WOLYNES || q->kind == AWSEM_Q_ONUCHIC_CUTOFF_GAUSS ||
                                                        q->kind == AWSEM_Q_ONUCHIC_SHADOW_GAUSS,
                  d_xbase, q->d_off, count, q->norm, q->d_out, st) ||
      hipMemcpyAsync(values, q->d_out, (size_t)count * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return qfail(q, AWSEM_E_DEVICE, "order parameter: kernel failed");
  return AWSEM_OK;
}

int aw_q_residues(const awsem_q *q) { return q ? q->n : -1; }

extern "C" {

// ComputeQWolynes (compute_q_wolynes.cpp:51-88, 152-167): n x n native
// distances, pairs |i - j| > sep, sigma = (1 + |i - j|)^sigma_exp,
// normalised by 2 / ((n - sep)(n - sep - 1))
int awsem_q_wolynes_create(const char *rnative_path, int n, int sep, double sigma_exp, awsem_q **out)
{
  if (!out || !rnative_path || n < 2) return AWSEM_E_ARG;
  awsem_q *q = new awsem_q();
  *out = q;
  q->kind = AWSEM_Q_WOLYNES;
  q->n = n;
  std::ifstream in(rnative_path);
  if (!in) return qfail(q, AWSEM_E_FILE, "Native distance file can't be read");
  std::vector<double> rn((size_t)n * n);
  for (size_t k = 0; k < rn.size(); ++k)
    if (!(in >> rn[k])) return qfail(q, AWSEM_E_FORMAT, "Native distance file: fewer than n x n values");
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (!(std::abs(j - i) > sep)) continue;
      double s = std::pow(1 + std::abs(i - j), sigma_exp);
      q->items.push_back(QItem{i, j, rn[(size_t)i * n + j], 1.0 / (2.0 * s * s)});
    }
  q->norm = 2.0 / (((double)n - (double)sep) * ((double)n - ((double)sep + 1.0)));
  return AWSEM_OK;
}

// ComputeQOnuchic::createContactArrays (compute_q_onuchic.cpp:150-209):
// cutoff kinds read "nAtoms" and the native CA coordinates, native pairs
// j - i > 3 closer than r_contact (the cutoff kind compares with factor^2
// times the native r^2); shadow kinds read "i j rn" lines (1-based, |i - j|
// > 3; a later line of a pair replaces it, every line counts in the
// normalisation).  The gauss kinds use sigma^2 = (1 + |i - j|)^0.3.
int awsem_q_onuchic_create(int kind, const char *path, int n, double r_contact, double factor, awsem_q **out)
{
  if (!out || !path || n < 1 || kind < AWSEM_Q_ONUCHIC_CUTOFF || kind > AWSEM_Q_ONUCHIC_SHADOW_GAUSS) return AWSEM_E_ARG;
  awsem_q *q = new awsem_q();
  *out = q;
  q->kind = kind;
  q->n = n;
  bool gauss = kind == AWSEM_Q_ONUCHIC_CUTOFF_GAUSS || kind == AWSEM_Q_ONUCHIC_SHADOW_GAUSS;
  auto sig = [](int i, int j) { return 1.0 / (2.0 * std::pow(1 + std::abs(i - j), 2.0 * 0.15)); };
  double count = 0.0;
  if (kind == AWSEM_Q_ONUCHIC_CUTOFF || kind == AWSEM_Q_ONUCHIC_CUTOFF_GAUSS) {
    FILE *f = std::fopen(path, "r");
    if (!f) return qfail(q, AWSEM_E_FILE, "Compute qonuchic: can't read native coordinates");
    int na = 0;
    if (std::fscanf(f, "%d", &na) != 1) na = -1;
    std::vector<double> x(3 * (size_t)std::max(na, 0));
    for (int i = 0; i < na; ++i)
      if (std::fscanf(f, "%lf %lf %lf", &x[3 * i], &x[3 * i + 1], &x[3 * i + 2]) != 3) {
        std::fclose(f);
        return qfail(q, AWSEM_E_FORMAT, "Compute qonuchic: native coordinates file format error");
      }
    std::fclose(f);
    if (na != n) return qfail(q, AWSEM_E_FORMAT, "Compute qonuchic: atom number mismatch");
    double rc2 = r_contact * r_contact;
    for (int i = 0; i < n; ++i)
      for (int j = i + 4; j < n; ++j) {
        double dx = x[3 * j] - x[3 * i], dy = x[3 * j + 1] - x[3 * i + 1], dz = x[3 * j + 2] - x[3 * i + 2];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 < rc2)) continue;
        count += 1.0;
        if (gauss) q->items.push_back(QItem{i, j, std::sqrt(r2), sig(i, j)});
        else q->items.push_back(QItem{i, j, r2 * factor * factor, 0.0});
      }
  } else {
    FILE *f = std::fopen(path, "r");
    if (!f) return qfail(q, AWSEM_E_FILE, "Compute qonuchic: can't read contact map file");
    std::map<std::pair<int, int>, double> last;
    int a, b;
    double rn;
    while (std::fscanf(f, "%d %d %lf", &a, &b, &rn) == 3) {
      --a;
      --b;
      if (a >= n || b >= n || a < 0 || b < 0 || std::abs(b - a) <= 3) {
        std::fclose(f);
        return qfail(q, AWSEM_E_FORMAT, "Compute qonuchic: wrong residue index in shadow contact map file");
      }
      last[{a, b}] = rn;
      count += 1.0;
    }
    std::fclose(f);
    // only i < j is visited (atoms in residue order); a (j, i) line counts
    // in the normalisation alone
    for (const auto &e : last) {
      int i = e.first.first, j = e.first.second;
      if (i >= j) continue;
      double r2 = e.second * e.second;
      if (gauss) q->items.push_back(QItem{i, j, std::sqrt(r2), sig(i, j)});
      else q->items.push_back(QItem{i, j, r2 * factor * factor, 0.0});
    }
  }
  q->norm = 1 / count;
  return AWSEM_OK;
}

int awsem_q_compute(awsem_q *q, const double *xca, double *value)
{
  if (!q || !xca || !value || q->n <= 0 || !q->err.empty()) return AWSEM_E_ARG;
  for (int k = 0; k < 3 * q->n; ++k)
    if (!std::isfinite(xca[k])) return qfail(q, AWSEM_E_NONFINITE, "non-finite coordinates");
  int rc = q_ready(q, 1);
  if (rc) return rc;
  if (hipMemcpy(q->d_x, xca, (size_t)q->n * 3 * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return qfail(q, AWSEM_E_DEVICE, "order parameter: copy failed");
  long long off = 0;
  return aw_q_eval(q, q->d_x, &off, 1, value, 0);
}

int awsem_q_last_error(const awsem_q *q, char *buf, size_t n)
{
  if (!q || !buf || n == 0) return AWSEM_E_ARG;
  std::snprintf(buf, n, "%s", q->err.c_str());
  return AWSEM_OK;
}

int awsem_q_destroy(awsem_q *q)
{
  delete q;
  return AWSEM_OK;
}

}  // extern "C"
