// awsem_frust.hip -- the frustratometers on the device: tertiary
// ([Tertiary_Frustratometer], fb.cpp:5057-5412, energies 5444-5548),
// n-mer ([Nmer_Frustratometer], fb.cpp:5600-6061) and fragment
// ([Fragment_Frustratometer], fb.cpp:4431-4832).
//
// A call at one coordinate set: densities rho_i (k_tf_rho, the well-0 sum
// of cWell::compute_ro, sml.h:630-650), then the host lists the rows and
// draws the decoys from the context's rand() stream in the reference's
// order, and the kernels evaluate every (row, entry) energy, entry 0 the
// native assignment and entries 1..m the decoys.  Tertiary contacts
// (k_tf_rows): the pair's water, burial and Debye-Hueckel terms, plus in
// mutational mode the environment sum over every other residue k, O(n) per
// entry, one workgroup per contact with the k-dependent, entry-independent
// pieces (the two wells' theta at r_xk, the density switch of k, the
// screened 1/r) computed once per chunk of k into LDS and shared by the
// contact's entries.  A single-residue entry depends on (residue, type)
// only: one n x 20 table (k_nm_single) serves every decoy.  FP64
// throughout.  The host reduces each row's decoys to mean and (population)
// standard deviation in the reference's order and writes the reference's
// output files.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/awsem.h"
#include "awsem_host.h"

// ---------------------------------------------------------------------------
// glibc's rand(): srand(seed) fills r[0] = seed, r[i] = 16807 r[i-1]
// mod (2^31 - 1) for i < 31, r[31..33] = r[0..2]; then r[i] = r[i-31] +
// r[i-3] (mod 2^32) and the k-th rand() is r[344 + k] >> 1 (the additive
// feedback generator of its random_r, TYPE_3).  The reference calls rand()
// without srand: seed 1.

AwFrustState::AwFrustState() { seed(1); }

AwFrustState::~AwFrustState()
{
  if (dat) std::fclose(dat);
  if (tcl) std::fclose(tcl);
  if (nm_dat) std::fclose(nm_dat);
  if (nm_tcl) std::fclose(nm_tcl);
  if (nm_trap) std::fclose(nm_trap);
  for (FILE *f : ff_out)
    if (f) std::fclose(f);
  for (void *p : ff_dev)
    if (p) (void)hipFree(p);
}

void AwFrustState::seed(unsigned s)
{
  long long r[34];
  r[0] = s ? s : 1;
  for (int i = 1; i < 31; ++i) r[i] = (16807LL * r[i - 1]) % 2147483647LL;
  for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
  for (int i = 0; i < 34; ++i) rng[i] = (uint32_t)r[i];
  rng_k = 0;  // the slot of r[34], 34 mod 34
  for (int i = 34; i < 344; ++i) next();
}

int AwFrustState::next()
{
  // the ring holds r[k-34 .. k-1] at slots k mod 34: r[k-31] at j + 3 and
  // r[k-3] at j + 31 (mod 34), j the slot of r[k]
  const int j = rng_k, a = j + 3 < 34 ? j + 3 : j + 3 - 34, b = j + 31 < 34 ? j + 31 : j + 31 - 34;
  uint32_t v = rng[a] + rng[b];
  rng[j] = v;
  rng_k = j + 1 < 34 ? j + 1 : 0;
  return (int)(v >> 1);
}

// ---------------------------------------------------------------------------
// device side

struct TfRow {
  int kind;        // 0 pair (configurational), 1 pair + environment (mutational)
  int a, b, cnt;   // residues; entries
  int nat, pad;    // 1: entry 0 is the native (its decoys follow), 0: every entry is a decoy
  long long off;   // first entry (in the energies, and for kind 0 in ent / geo)
  long long doff;  // kind 1: the decoys' residue pairs at draws[doff + 2 d], [doff + 2 d + 1]
  double r, ra, rb;  // pair distance and densities (kind 1; kind 0 per entry)
};
struct TfEnt {
  int ar, at, br, bt;  // residue index (phosphorylation, minimum separation) and type of each side
};
struct TfGeo {
  double r, ra, rb;    // kind 0: the entry's distance and densities
};

__device__ __forceinline__ const double *tf_site(const double *x, const int *rinfo, int n, int i)
{
  return x + (((rinfo[i] & RI_GLY) ? 0 : 1) * n + i) * 3;
}

__device__ __forceinline__ double tf_dist(const double *x, const int *rinfo, int n, int i, int j)
{
  const double *p = tf_site(x, rinfo, n, i), *q = tf_site(x, rinfo, n, j);
  double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// the well-0 density of cWell::compute_ro (sml.h:630-650, 563-588): pairs
// of another chain or |i - j| > 1, theta 0 outside (rmin_theta, rmax_theta).
// One wave per residue: the lanes stride over j and the partial sums meet
// in a fixed butterfly, so the value does not depend on the launch.
__global__ void __launch_bounds__(256) k_tf_rho(const FFParams *__restrict__ P, const int *__restrict__ rinfo,
                                               const double *__restrict__ x, int n, double *__restrict__ rho)
{
  const int i = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (i >= n) return;  // a whole wave leaves together
  const double *si = tf_site(x, rinfo, n, i);
  const double k = P->water_kappa, r0 = P->well_r_min[0], r1 = P->well_r_max[0];
  const double lo2 = P->w_rmin_th_sq[0], hi2 = P->w_rmax_th_sq[0];
  const int ci = rinfo[i] >> RI_CHAIN_SHIFT;
  double s = 0.0;
  for (int j = lane; j < n; j += 64) {
    if (!((rinfo[j] >> RI_CHAIN_SHIFT) != ci || abs(j - i) > 1)) continue;
    const double *sj = tf_site(x, rinfo, n, j);
    double dx = si[0] - sj[0], dy = si[1] - sj[1], dz = si[2] - sj[2];
    double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < lo2 || r2 > hi2) continue;
    double r = sqrt(r2);
    s += 0.25 * (1.0 + tanh(k * (r - r0))) * (1.0 + tanh(k * (r1 - r)));
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) rho[i] = s;
}

__device__ __forceinline__ const double (*tf_wg(const FFParams *P, int ri, int rj))[20][20][2]
{
  return ((ri | rj) & RI_PHOS) ? P->phosph_gamma : P->water_gamma;
}

// compute_water_energy (fb.cpp:5444-5476) from the two wells' thetas and
// the density switches h = 1 - tanh(kappa_sigma (rho - rho0))
__device__ __forceinline__ double tf_water(const FFParams *P, const int *rinfo, int ir, int jr, int it, int jt,
                                           double td, double tm, double hi, double hj)
{
  const double(*g)[20][20][2] = tf_wg(P, rinfo[ir], rinfo[jr]);
  double sw = 0.25 * hi * hj, sp = 1.0 - sw;
  double sgd = (g[0][it][jt][0] + g[0][it][jt][1]) / 2, sgm = sp * g[1][it][jt][0] + sw * g[1][it][jt][1];
  return -(sgd * td + sgm * tm);
}

__device__ __forceinline__ void tf_thetas(const FFParams *P, double r, double &td, double &tm)
{
  double k = P->water_kappa;
  td = 0.25 * (1.0 + tanh(k * (r - P->well_r_min[0]))) * (1.0 + tanh(k * (P->well_r_max[0] - r)));
  tm = 0.25 * (1.0 + tanh(k * (r - P->well_r_min[1]))) * (1.0 + tanh(k * (P->well_r_max[1] - r)));
}

// compute_burial_energy (fb.cpp:5478-5500); a phosphorylated residue reads type 6
__device__ __forceinline__ double tf_burial(const FFParams *P, const int *rinfo, int ir, int it, double rho)
{
  int t = (rinfo[ir] & RI_PHOS) ? 6 : it;
  double kb = P->burial_kappa, e = 0.0;
  for (int b = 0; b < 3; ++b)
    e += -0.5 * P->k_burial * P->burial_gamma[t][b] * (tanh(kb * (rho - P->burial_ro_min[b])) + tanh(kb * (P->burial_ro_max[b] - rho)));
  return e;
}

// charge of a residue type (fb.cpp:5510-5531): R, K +1; D, E -1
__device__ __forceinline__ double tf_charge(int t) { return (t == 1 || t == 11) ? 1.0 : ((t == 3 || t == 6) ? -1.0 : 0.0); }

// compute_electrostatic_energy (fb.cpp:5502-5548) with the screened 1/r
// factor s = exp(-k_s r / lambda) / r given
__device__ __forceinline__ double tf_elec(const FFParams *P, int ir, int jr, int it, int jt, double s)
{
  if (abs(ir - jr) < P->dh_min_sep) return 0.0;
  double qi = tf_charge(it), qj = tf_charge(jt);
  if (qi == 0.0 || qj == 0.0) return 0.0;
  double k = (qi > 0 && qj > 0) ? P->dh_k[0] : ((qi < 0 && qj < 0) ? P->dh_k[1] : P->dh_k[2]);
  return P->epsilon * k * qi * qj * s;
}

#define TF_CH 512  // residues k per LDS chunk

// entry e of a row: the residues whose types (and indices) it assigns
__device__ __forceinline__ TfEnt tf_entry(const TfRow &R, int e, const TfEnt *__restrict__ ent,
                                          const int *__restrict__ draws, const int *__restrict__ rinfo)
{
  if (R.kind == 0) return ent[R.off + e];
  int a = R.a, b = R.b;
  if (e > 0) {
    a = draws[R.doff + 2 * (e - 1)];
    b = draws[R.doff + 2 * (e - 1) + 1];
  }
  return TfEnt{a, rinfo[a] & RI_TYPE_MASK, b, rinfo[b] & RI_TYPE_MASK};
}

// a workgroup's sum of one value per thread (fixed tree: deterministic)
__device__ __forceinline__ double tf_block_sum(double v, double *red)
{
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  double s = red[0];
  __syncthreads();
  return s;
}

// the mutational environment: sides a and b over k != a, b (water within
// the cutoff, Debye-Hueckel always, fb.cpp:5211-5239), added to the row's
// entries; the k-dependent factors staged in LDS a chunk of TF_CH at a time
__device__ __forceinline__ void tf_env_pass(const FFParams *__restrict__ P, const int *__restrict__ rinfo,
                                            const double *__restrict__ x, int n, const double *__restrict__ rho,
                                            const TfRow &R, const TfEnt *__restrict__ ent,
                                            const int *__restrict__ draws, double *__restrict__ out,
                                            double (*c_td)[TF_CH], double (*c_tm)[TF_CH], double (*c_es)[TF_CH],
                                            double *c_hk, int (*c_on)[TF_CH], int *c_tk)
{
  const bool dh = P->dh_flag != 0;
  const double ks = P->water_kappa_sigma, th = P->treshold;
  const int ns = 2;
  const int side[2] = {R.a, R.b};
  const double hs[2] = {1.0 - tanh(ks * (R.ra - th)), 1.0 - tanh(ks * (R.rb - th))};
  const double cut = P->tf_cutoff;
  for (int k0 = 0; k0 < n; k0 += TF_CH) {
    int kn = min(TF_CH, n - k0);
    __syncthreads();
    for (int q = threadIdx.x; q < kn; q += blockDim.x) {
      int k = k0 + q;
      c_hk[q] = 1.0 - tanh(ks * (rho[k] - th));
      c_tk[q] = rinfo[k] & RI_TYPE_MASK;
      for (int s = 0; s < ns; ++s) {
        int a = side[s];
        bool skip = k == R.a || k == R.b;
        double r = tf_dist(x, rinfo, n, a, k);
        bool wat = !skip && r < cut;
        double td = 0.0, tm = 0.0;
        if (wat) tf_thetas(P, r, td, tm);
        c_td[s][q] = td;
        c_tm[s][q] = tm;
        c_on[s][q] = (wat ? 1 : 0) | (!skip && dh ? 2 : 0);
        c_es[s][q] = (!skip && dh) ? exp(-P->dh_kscreen * r / P->dh_length) / r : 0.0;
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < R.cnt; e += blockDim.x) {
      TfEnt E = tf_entry(R, e, ent, draws, rinfo);
      int er[2] = {E.ar, E.br}, et[2] = {E.at, E.bt};
      double acc = 0.0;
      for (int q = 0; q < kn; ++q) {
        int k = k0 + q, tk = c_tk[q];
        for (int s = 0; s < ns; ++s) {
          int on = c_on[s][q];
          if (on & 1) acc += tf_water(P, rinfo, er[s], k, et[s], tk, c_td[s][q], c_tm[s][q], hs[s], c_hk[q]);
          if (on & 2) acc += tf_elec(P, er[s], k, et[s], tk, c_es[s][q]);
        }
      }
      out[R.off + e] += acc;
    }
  }
}

__global__ void __launch_bounds__(256) k_tf_rows(const FFParams *__restrict__ P, const int *__restrict__ rinfo,
                                                const double *__restrict__ x, int n, const double *__restrict__ rho,
                                                const TfRow *__restrict__ rows, const TfEnt *__restrict__ ent,
                                                const TfGeo *__restrict__ geo, const int *__restrict__ draws,
                                                double *__restrict__ out, double *__restrict__ stat)
{
  __shared__ double c_td[2][TF_CH], c_tm[2][TF_CH], c_es[2][TF_CH], c_hk[TF_CH];
  __shared__ int c_on[2][TF_CH], c_tk[TF_CH];
  __shared__ double red[256];
  const TfRow R = rows[blockIdx.x];
  const bool dh = P->dh_flag != 0;
  const double ks = P->water_kappa_sigma, th = P->treshold;
  // base: the pair's water, burial and electrostatics
  for (int e = threadIdx.x; e < R.cnt; e += blockDim.x) {
    TfEnt E = tf_entry(R, e, ent, draws, rinfo);
    double r = R.r, ra = R.ra, rb = R.rb;
    if (R.kind == 0) {
      TfGeo G = geo[R.off + e];
      r = G.r; ra = G.ra; rb = G.rb;
    }
    double td, tm;
    tf_thetas(P, r, td, tm);
    double v = tf_water(P, rinfo, E.ar, E.br, E.at, E.bt, td, tm, 1.0 - tanh(ks * (ra - th)), 1.0 - tanh(ks * (rb - th)));
    v += tf_burial(P, rinfo, E.ar, E.at, ra) + tf_burial(P, rinfo, E.br, E.bt, rb);
    if (dh) v += tf_elec(P, E.ar, E.br, E.at, E.bt, exp(-P->dh_kscreen * r / P->dh_length) / r);
    out[R.off + e] = v;
  }
  if (R.kind == 1) tf_env_pass(P, rinfo, x, n, rho, R, ent, draws, out, c_td, c_tm, c_es, c_hk, c_on, c_tk);
  // the row's decoy statistics (compute_array_mean / _std, fb.cpp:5413-5442:
  // population standard deviation)
  __syncthreads();
  const int m = R.cnt - R.nat;
  double sum = 0.0;
  for (int e = R.nat + threadIdx.x; e < R.cnt; e += blockDim.x) sum += out[R.off + e];
  const double mean = m > 0 ? tf_block_sum(sum, red) / m : 0.0;
  double sq = 0.0;
  for (int e = R.nat + threadIdx.x; e < R.cnt; e += blockDim.x) {
    double d = out[R.off + e] - mean;
    sq += d * d;
  }
  const double var = m > 0 ? tf_block_sum(sq, red) / m : 0.0;
  if (threadIdx.x == 0) {
    stat[3 * (size_t)blockIdx.x] = R.nat ? out[R.off] : 0.0;
    stat[3 * (size_t)blockIdx.x + 1] = mean;
    stat[3 * (size_t)blockIdx.x + 2] = sqrt(var);
  }
}

// ---------------------------------------------------------------------------
// the n-mer frustratometer (fb.cpp:5600-6061)

#define NM_MAX 32  // largest n-mer size the kernels take

// contacts between the n-mers at i and j (compute_nmer_contacts,
// fb.cpp:5895-5919: the sequence-separation test only, no chain test) for
// the start pairs i < j < n - s with j - i >= s, rows i0 <= i < i1 (entry
// (i - i0) ns + j); -1 for the other pairs
__global__ void __launch_bounds__(256) k_nm_contacts(const FFParams *__restrict__ P, const int *__restrict__ rinfo,
                                                    const double *__restrict__ x, int n, int s, int i0, int i1,
                                                    int *__restrict__ cnt)
{
  const int ns = n - s;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)(i1 - i0) * ns) return;
  int i = i0 + (int)(t / ns), j = (int)(t % ns);
  if (!(j > i && j - i >= s)) {
    cnt[t] = -1;
    return;
  }
  int c = 0;
  const double cut = P->nf_cutoff;
  for (int a = i; a < i + s; ++a)
    for (int b = j; b < j + s; ++b)
      if (abs(a - b) >= P->contact_cutoff && tf_dist(x, rinfo, n, a, b) < cut) c++;
  cnt[t] = c;
}

struct NmRow {
  int a, b, cnt, ncont;  // the two n-mers' first residues; entries; contacts
  long long off;         // first entry
  long long doff;        // the decoys' starts at draws[doff + 2 d], [doff + 2 d + 1]
  int m, nt;             // decoys, trap entries (0 or 4)
};
// the types of n-mer a are those of residues sa + da*o, of n-mer b those
// of sb + db*o; swap: the reference's order is b's burials first and b
// outer in the water loop (compute_nmer_traps(j, i), fb.cpp:5622)
struct NmEnt {
  int sa, sb;
  int da, db, swap, pad;
};

// entry e of a row: 0 the native, 1..m the decoys, then the four trap
// entries: the second n-mer's types from residues o (forward) or s - o
// (backward), for (i, j) and, in the reference's order, (j, i)
// (fb.cpp:5829, 5839: the trap types are read without k)
__device__ __forceinline__ NmEnt nm_entry(const NmRow &R, int e, const int *__restrict__ draws, int s)
{
  if (e == 0) return NmEnt{R.a, R.b, 1, 1, 0, 0};
  if (e <= R.m) return NmEnt{draws[R.doff + 2 * (e - 1)], draws[R.doff + 2 * (e - 1) + 1], 1, 1, 0, 0};
  switch (e - 1 - R.m) {
    case 0: return NmEnt{R.a, 0, 1, 1, 0, 0};
    case 1: return NmEnt{R.a, s, 1, -1, 0, 0};
    case 2: return NmEnt{0, R.b, 1, 1, 1, 0};
    default: return NmEnt{s, R.b, -1, 1, 1, 0};
  }
}

// every (row, entry) energy of the pairwise mode: burials of both n-mers
// and the water of their s*s cross pairs (compute_nmer_native_ixn,
// fb.cpp:5947-5988, the decoys 6028-6056, the traps 5822-5853).  One
// workgroup per row; the pair geometry (both wells' theta), the density
// switches and every member's burial for each of the 20 types are staged
// in LDS once and shared by the row's entries.
__global__ void __launch_bounds__(256) k_nm_pairs(const FFParams *__restrict__ P, const int *__restrict__ rinfo,
                                                 const double *__restrict__ x, int n, int s,
                                                 const double *__restrict__ rho, const NmRow *__restrict__ rows,
                                                 const int *__restrict__ draws, double *__restrict__ out,
                                                 double *__restrict__ stat)
{
  __shared__ double l_td[NM_MAX * NM_MAX], l_tm[NM_MAX * NM_MAX];
  __shared__ double red[256];
  __shared__ double l_h[2][NM_MAX];
  __shared__ double l_bur[2][NM_MAX][20];
  const NmRow R = rows[blockIdx.x];
  const double ks = P->water_kappa_sigma, th = P->treshold;
  for (int q = threadIdx.x; q < s * s; q += blockDim.x) {
    int o = q / s, p = q % s;
    double td, tm;
    tf_thetas(P, tf_dist(x, rinfo, n, R.a + o, R.b + p), td, tm);
    l_td[q] = td;
    l_tm[q] = tm;
  }
  for (int q = threadIdx.x; q < 2 * s * 20; q += blockDim.x) {
    int side = q / (s * 20), o = (q / 20) % s, t = q % 20;
    int r = (side ? R.b : R.a) + o;
    l_bur[side][o][t] = tf_burial(P, rinfo, r, t, rho[r]);
  }
  for (int q = threadIdx.x; q < 2 * s; q += blockDim.x) {
    int side = q / s, o = q % s;
    l_h[side][o] = 1.0 - tanh(ks * (rho[(side ? R.b : R.a) + o] - th));
  }
  __syncthreads();
  for (int e = threadIdx.x; e < R.cnt; e += blockDim.x) {
    const NmEnt E = nm_entry(R, e, draws, s);
    double v = 0.0;
    if (!E.swap) {
      for (int o = 0; o < s; ++o) v += l_bur[0][o][rinfo[E.sa + E.da * o] & RI_TYPE_MASK];
      for (int p = 0; p < s; ++p) v += l_bur[1][p][rinfo[E.sb + E.db * p] & RI_TYPE_MASK];
      for (int o = 0; o < s; ++o) {
        int ta = rinfo[E.sa + E.da * o] & RI_TYPE_MASK;
        for (int p = 0; p < s; ++p) {
          int tb = rinfo[E.sb + E.db * p] & RI_TYPE_MASK;
          v += tf_water(P, rinfo, R.a + o, R.b + p, ta, tb, l_td[o * s + p], l_tm[o * s + p], l_h[0][o], l_h[1][p]);
        }
      }
    } else {
      for (int p = 0; p < s; ++p) v += l_bur[1][p][rinfo[E.sb + E.db * p] & RI_TYPE_MASK];
      for (int o = 0; o < s; ++o) v += l_bur[0][o][rinfo[E.sa + E.da * o] & RI_TYPE_MASK];
      for (int p = 0; p < s; ++p) {
        int tb = rinfo[E.sb + E.db * p] & RI_TYPE_MASK;
        for (int o = 0; o < s; ++o) {
          int ta = rinfo[E.sa + E.da * o] & RI_TYPE_MASK;
          v += tf_water(P, rinfo, R.b + p, R.a + o, tb, ta, l_td[o * s + p], l_tm[o * s + p], l_h[1][p], l_h[0][o]);
        }
      }
    }
    out[R.off + e] = v;
  }
  // the row's decoy mean and (population) standard deviation
  // (fb.cpp:6058-6059), its native and trap energies: 7 values per row
  __syncthreads();
  double sum = 0.0;
  for (int e = 1 + threadIdx.x; e <= R.m; e += blockDim.x) sum += out[R.off + e];
  const double mean = tf_block_sum(sum, red) / R.m;
  double sq = 0.0;
  for (int e = 1 + threadIdx.x; e <= R.m; e += blockDim.x) {
    double d = out[R.off + e] - mean;
    sq += d * d;
  }
  const double var = tf_block_sum(sq, red) / R.m;
  if (threadIdx.x < 7) {
    const int q = threadIdx.x;
    double v = q == 0 ? out[R.off] : (q == 1 ? mean : (q == 2 ? sqrt(var) : 0.0));
    if (q >= 3 && q - 3 < R.nt) v = out[R.off + 1 + R.m + (q - 3)];
    stat[7 * (size_t)blockIdx.x + q] = v;
  }
}

// the energy of residue j with type t for every (j, t)
// (compute_singleresidue_native_ixn, fb.cpp:5346-5393: its burial, its
// water with the other residues within the cutoff and the contact
// separation or another chain, its Debye-Hueckel with every other residue;
// with lower_only (nmercalc) the residues k < j only), summed as the
// reference sums: water + burial + electrostatics.  A single-residue row's
// entry is then one table value, a single n-mer's a sum of s of them.
__global__ void __launch_bounds__(256) k_nm_single(const FFParams *__restrict__ P, const int *__restrict__ rinfo,
                                                  const double *__restrict__ x, int n,
                                                  const double *__restrict__ rho, double cut, int lower_only,
                                                  double *__restrict__ e1)
{
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * 20) return;
  const int j = q / 20, t = q % 20;
  const double ks = P->water_kappa_sigma, th = P->treshold;
  const bool dh = P->dh_flag != 0;
  const double hj = 1.0 - tanh(ks * (rho[j] - th));
  const int cj = rinfo[j] >> RI_CHAIN_SHIFT;
  double w = 0.0, el = 0.0;
  const int kend = lower_only ? j : n;
  for (int k = 0; k < kend; ++k) {
    if (k == j) continue;
    double r = tf_dist(x, rinfo, n, j, k);
    int kt = rinfo[k] & RI_TYPE_MASK;
    if (r < cut && (abs(j - k) >= P->contact_cutoff || (rinfo[k] >> RI_CHAIN_SHIFT) != cj)) {
      double td, tm;
      tf_thetas(P, r, td, tm);
      w += tf_water(P, rinfo, j, k, t, kt, td, tm, hj, 1.0 - tanh(ks * (rho[k] - th)));
    }
    if (dh) el += tf_elec(P, j, k, t, kt, exp(-P->dh_kscreen * r / P->dh_length) / r);
  }
  e1[q] = w + tf_burial(P, rinfo, j, t, rho[j]) + el;
}

// entry e of row e / (m + 1): the n-mer at row's start with the types of
// residues src[e] + o (compute_singlenmer_native_ixn / decoy_ixns,
// fb.cpp:5708-5768)
__global__ void __launch_bounds__(256) k_nm_single_sum(const int *__restrict__ rinfo, const double *__restrict__ e1,
                                                      int s, int m, long long n_ent, const int *__restrict__ src,
                                                      double *__restrict__ out)
{
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_ent) return;
  const int i = (int)(e / (m + 1)), sr = src[e];
  double v = 0.0;
  for (int o = 0; o < s; ++o) v += e1[(i + o) * 20 + (rinfo[sr + o] & RI_TYPE_MASK)];
  out[e] = v;
}

// ---------------------------------------------------------------------------
// the fragment frustratometer (fb.cpp:4431-4832)

// one (memory, relative pair) of a library: the pair (di, dj) of the
// memory's residues that compute_decoy_memory_potential visits (residue
// pos + di of the map, fb.cpp:4478-4496, partner pos + dj), the memory's four
// reference distances (CA-CA, CA(i)-CB(j), CB(i)-CA(j), CB-CB, fm.cpp:190-202)
// and its glycine flags (bit 0 at di, bit 1 at dj)
struct FfItem {
  int k, di, dj, gly;
  double rf[4];
};

// the Gaussians of one item: target residues i, j with distance r[q] per
// atom pair q; the reference adds each to E[i] and E[j] (fb.cpp:4497-4561),
// here their sum is returned and added to both once
__device__ __forceinline__ double ff_pair(const FFParams *__restrict__ P, const int *__restrict__ rinfo,
                                          const double *__restrict__ gsep, const FfItem &I, double w, int i, int j,
                                          const double r[4])
{
  const int ti = rinfo[i] & RI_TYPE_MASK, tj = rinfo[j] & RI_TYPE_MASK, d = j - i;
  double g = gsep[(size_t)d * 400 + ti * 20 + tj];
  if (P->ff_seqsep) g *= pow((double)(d - P->fm_min_sep + 1), -P->ff_seqsep_gamma);
  const double ekwg = P->epsilon * P->k_frag_mem * w * g;
  double s2 = pow((double)d, 2 * P->fm_sigma_exp);
  s2 = s2 * P->ff_ww * P->ff_ww;
  const bool gi = (rinfo[i] & RI_GLY) || (I.gly & 1), gj = (rinfo[j] & RI_GLY) || (I.gly & 2);
  double v = 0.0;
  for (int q = 0; q < 4; ++q) {
    if ((q >= 2 && gi) || ((q & 1) && gj)) continue;
    double dr = r[q] - I.rf[q];
    v += -ekwg * exp(-(dr * dr) / (2 * s2));
  }
  return v;
}

__device__ __forceinline__ double ff_r(const double *a, const double *b)
{
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// a library at the current coordinates, one calculation per blockIdx.y
// with the memories at pos[calc * nm + k]: E[calc * n + i]
// (compute_decoy_memory_potential, fb.cpp:4431-4563)
__global__ void __launch_bounds__(256) k_ff_mem(const FFParams *__restrict__ P, const int *__restrict__ rinfo,
                                               const double *__restrict__ x, int n, const FfItem *__restrict__ items,
                                               int n_items, const int *__restrict__ pos, const double *__restrict__ w,
                                               int nm, const double *__restrict__ gsep, double *__restrict__ E,
                                               int *__restrict__ err)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_items) return;
  const int c = blockIdx.y;
  const FfItem I = items[t];
  const int p = pos[(size_t)c * nm + I.k], i = p + I.di, j = p + I.dj;
  if ((rinfo[i] >> RI_CHAIN_SHIFT) != (rinfo[j] >> RI_CHAIN_SHIFT)) {
    atomicOr(err, 1);
    return;
  }
  const double *ci = x + (size_t)i * 3, *cj = x + (size_t)j * 3;
  const double *bi = x + ((size_t)n + i) * 3, *bj = x + ((size_t)n + j) * 3;
  const double r[4] = {ff_r(ci, cj), ff_r(ci, bj), ff_r(bi, cj), ff_r(bi, bj)};
  const double v = ff_pair(P, rinfo, gsep, I, w[I.k], i, j, r);
  atomicAdd(&E[(size_t)c * n + i], v);
  atomicAdd(&E[(size_t)c * n + j], v);
}

// the same with the calculation's n energies accumulated in LDS by each
// workgroup over FF_CHUNK items and added to E once per workgroup: the
// items of one calculation hit few residues, so global atomics would
// serialise on them (n * 8 bytes of LDS, n <= FF_LDS_MAX)
#define FF_CHUNK 2048
#define FF_LDS_MAX 8192
__global__ void __launch_bounds__(256) k_ff_mem_lds(const FFParams *__restrict__ P, const int *__restrict__ rinfo,
                                                   const double *__restrict__ x, int n,
                                                   const FfItem *__restrict__ items, int n_items,
                                                   const int *__restrict__ pos, const double *__restrict__ w, int nm,
                                                   const double *__restrict__ gsep, double *__restrict__ E,
                                                   int *__restrict__ err)
{
  extern __shared__ double e_l[];
  const int c = blockIdx.y;
  for (int q = threadIdx.x; q < n; q += blockDim.x) e_l[q] = 0.0;
  __syncthreads();
  const int t1 = min(n_items, (int)(blockIdx.x + 1) * FF_CHUNK);
  for (int t = blockIdx.x * FF_CHUNK + threadIdx.x; t < t1; t += blockDim.x) {
    const FfItem I = items[t];
    const int p = pos[(size_t)c * nm + I.k], i = p + I.di, j = p + I.dj;
    if ((rinfo[i] >> RI_CHAIN_SHIFT) != (rinfo[j] >> RI_CHAIN_SHIFT)) {
      atomicOr(err, 1);
      continue;
    }
    const double *ci = x + (size_t)i * 3, *cj = x + (size_t)j * 3;
    const double *bi = x + ((size_t)n + i) * 3, *bj = x + ((size_t)n + j) * 3;
    const double r[4] = {ff_r(ci, cj), ff_r(ci, bj), ff_r(bi, cj), ff_r(bi, bj)};
    const double v = ff_pair(P, rinfo, gsep, I, w[I.k], i, j, r);
    atomicAdd(&e_l[i], v);
    atomicAdd(&e_l[j], v);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += blockDim.x)
    if (e_l[q] != 0.0) atomicAdd(&E[(size_t)c * n + q], e_l[q]);
}

// read mode's decoy energies: the memories at their positions compared with
// decoy structure blockIdx.y (compute_generated_decoy_energies,
// fb.cpp:4691-4784): the decoy's distance where it covers both residues,
// else 0 (its Rf returns 0 there, unchecked)
__global__ void __launch_bounds__(256) k_ff_gen(const FFParams *__restrict__ P, const int *__restrict__ rinfo, int n,
                                               const FfItem *__restrict__ items, int n_items,
                                               const int *__restrict__ pos, const double *__restrict__ w,
                                               const double *__restrict__ gsep, const double *__restrict__ dxyz,
                                               const int *__restrict__ dinfo, double *__restrict__ G,
                                               int *__restrict__ err)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_items) return;
  const int d = blockIdx.y;
  const FfItem I = items[t];
  const int p = pos[I.k], i = p + I.di, j = p + I.dj;
  if ((rinfo[i] >> RI_CHAIN_SHIFT) != (rinfo[j] >> RI_CHAIN_SHIFT)) {
    atomicOr(err, 1);
    return;
  }
  // decoy d: first residue dpos, length dlen, its CA then CB coordinates at dxyz + off
  const int dpos = dinfo[3 * d], dlen = dinfo[3 * d + 1], off = dinfo[3 * d + 2];
  double r[4] = {0.0, 0.0, 0.0, 0.0};
  const int a = i - dpos, b = j - dpos;
  if (a >= 0 && a < dlen && b >= 0 && b < dlen) {
    const double *ca = dxyz + off, *cb = dxyz + off + 3 * dlen;
    r[0] = ff_r(ca + 3 * a, ca + 3 * b);
    r[1] = ff_r(ca + 3 * a, cb + 3 * b);
    r[2] = ff_r(cb + 3 * a, ca + 3 * b);
    r[3] = ff_r(cb + 3 * a, cb + 3 * b);
  }
  const double v = ff_pair(P, rinfo, gsep, I, w[I.k], i, j, r);
  atomicAdd(&G[(size_t)d * n + i], v);
  atomicAdd(&G[(size_t)d * n + j], v);
}

// ---------------------------------------------------------------------------
// host side

#define TF_HIPCHK(expr)                                                                        \
  do {                                                                                         \
    if ((expr) != hipSuccess) return AwStatus::err(AWSEM_E_DEVICE, std::string("frustratometer: ") + #expr); \
  } while (0)

namespace {

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  template <class T>
  AwStatus put(const std::vector<T> &v, hipStream_t st)
  {
    size_t b = std::max<size_t>(v.size(), 1) * sizeof(T);
    if (hipMalloc(&p, b) != hipSuccess) return AwStatus::err(AWSEM_E_NOMEM, "frustratometer: device allocation");
    if (!v.empty() && hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st) != hipSuccess)
      return AwStatus::err(AWSEM_E_DEVICE, "frustratometer: copy");
    return AwStatus();
  }
};

// a vector copied to device memory that lives with the context's
// frustratometer state (freed with it)
template <class T>
AwStatus put_static(const std::vector<T> &v, void *&p)
{
  if (p) (void)hipFree(p);
  p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(v.size(), 1) * sizeof(T)) != hipSuccess)
    return AwStatus::err(AWSEM_E_NOMEM, "frustratometer: device allocation");
  if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return AwStatus::err(AWSEM_E_DEVICE, "frustratometer: copy");
  return AwStatus();
}

// site (CB; CA for glycine) of residue i in the residue layout
inline const double *site(const AwSystem &S, const double *x, int i)
{
  return x + ((S.se[i] == 'G' ? 0 : 1) * (size_t)S.n + i) * 3;
}
inline double dist(const AwSystem &S, const double *x, int i, int j)
{
  const double *p = site(S, x, i), *q = site(S, x, j);
  double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

void stats(const double *e, int m, double *st)
{
  double mean = 0.0, var = 0.0;
  for (int k = 0; k < m; ++k) mean += e[k];
  mean /= (double)m;
  for (int k = 0; k < m; ++k) var += (e[k] - mean) * (e[k] - mean);
  st[0] = mean;
  st[1] = std::sqrt(var / (double)m);
}

}  // namespace

// one call: rows of 11 doubles (i, j, i_chain, j_chain, r_ij, rho_i, rho_j,
// native, decoy mean, decoy std, index; 1-based; j = j_chain = 0 for the
// single-residue mode)
static AwStatus tert_frust(awsem_ctx *c, const double *xres, std::vector<double> &table)
{
  const AwSystem &S = *aw_ctx_system(c);
  const FFParams &P = S.P;
  if (!P.tf_flag) return AwStatus::err(AWSEM_E_STATE, "no [Tertiary_Frustratometer] section in the coeff file");
  if (P.tf_ndecoys < 1) return AwStatus::err(AWSEM_E_FORMAT, "Tertiary_Frustratometer: no decoys");
  AwCtxDev D;
  AwStatus s = aw_ctx_device(c, xres, D);
  if (!s.ok()) return s;
  hipStream_t st = (hipStream_t)D.stream;
  int n = S.n, m = P.tf_ndecoys;
  AwFrustState &F = aw_ctx_frust(c);
  // densities on the device
  double *d_rho = nullptr;
  TF_HIPCHK(hipMalloc((void **)&d_rho, (size_t)n * sizeof(double)));
  DevBuf rho_buf;
  rho_buf.p = d_rho;
  hipLaunchKernelGGL(k_tf_rho, dim3((n + 3) / 4), dim3(256), 0, st, D.params, D.rinfo, D.x, n, d_rho);
  std::vector<double> rho(n);
  TF_HIPCHK(hipMemcpyAsync(rho.data(), d_rho, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
  TF_HIPCHK(hipStreamSynchronize(st));
  auto type = [&](int i) { return S.rinfo[i] & RI_TYPE_MASK; };
  auto chain = [&](int i) { return S.rinfo[i] >> RI_CHAIN_SHIFT; };
  double cut = P.tf_cutoff;
  std::vector<TfRow> rows;
  std::vector<TfEnt> ent;
  std::vector<TfGeo> geo;
  std::vector<int> draws;
  struct Pair { int i, j; double r; };
  std::vector<Pair> contacts;
  long long off = 0;
  // configurational: the decoy row, once per simulation (fb.cpp:5263-5290)
  int conf_row = -1;
  if (P.tf_mode != 2) {
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) {
        double r = dist(S, xres, i, j);
        if (r < cut && (std::abs(i - j) >= P.contact_cutoff || chain(i) != chain(j))) contacts.push_back(Pair{i, j, r});
      }
  }
  if (P.tf_mode == 0) {
    for (const Pair &q : contacts) {
      rows.push_back(TfRow{0, q.i, q.j, 1, 1, 0, off, 0, q.r, rho[q.i], rho[q.j]});
      ent.push_back(TfEnt{q.i, type(q.i), q.j, type(q.j)});
      geo.push_back(TfGeo{q.r, rho[q.i], rho[q.j]});
      off += 1;
    }
    if (!F.done && !contacts.empty()) {
      conf_row = (int)rows.size();
      rows.push_back(TfRow{0, 0, 0, m, 0, 0, off, 0, 0.0, 0.0, 0.0});
      for (int d = 0; d < m; ++d) {
        int a = F.next() % n, b = F.next() % n;
        double r = dist(S, xres, a, b);
        while (r > cut || a == b) {
          a = F.next() % n;
          b = F.next() % n;
          r = dist(S, xres, a, b);
        }
        int a2 = F.next() % n, b2 = F.next() % n;
        int a3 = F.next() % n, b3 = F.next() % n;
        ent.push_back(TfEnt{a3, type(a3), b3, type(b3)});
        geo.push_back(TfGeo{r, rho[a2], rho[b2]});
      }
      off += m;
    }
  } else if (P.tf_mode == 1) {
    // the decoys' residue pairs only: the kernel reads their types
    draws.reserve(contacts.size() * 2 * (size_t)m);
    for (const Pair &q : contacts) {
      rows.push_back(TfRow{1, q.i, q.j, m + 1, 1, 0, off, (long long)draws.size(), q.r, rho[q.i], rho[q.j]});
      for (int d = 0; d < m; ++d) {
        int a = F.next() % n, b = F.next() % n;
        draws.push_back(a);
        draws.push_back(b);
      }
      off += m + 1;
    }
  } else {
    for (int i = 0; i < n; ++i) {
      rows.push_back(TfRow{2, i, -1, m + 1, 1, 0, off, 0, 0.0, rho[i], 0.0});
      ent.push_back(TfEnt{i, type(i), 0, 0});
      for (int d = 0; d < m; ++d) ent.push_back(TfEnt{i, type(F.next() % n), 0, 0});
      off += m + 1;
    }
  }
  std::vector<double> stat(rows.size() * 3);  // per row: native, decoy mean, decoy std
  if (P.tf_mode == 2) {
    // a single-residue entry depends on (residue, type) only: one table of
    // n x 20 energies (compute_singleresidue_native_ixn, fb.cpp:5346-5393)
    DevBuf b_t;
    TF_HIPCHK(hipMalloc(&b_t.p, (size_t)n * 20 * sizeof(double)));
    hipLaunchKernelGGL(k_nm_single, dim3((n * 20 + 255) / 256), dim3(256), 0, st, D.params, D.rinfo, D.x, n, d_rho,
                       cut, 0, (double *)b_t.p);
    TF_HIPCHK(hipGetLastError());
    std::vector<double> T((size_t)n * 20);
    TF_HIPCHK(hipMemcpyAsync(T.data(), b_t.p, T.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    TF_HIPCHK(hipStreamSynchronize(st));
    std::vector<double> e((size_t)off);
    for (size_t k = 0; k < e.size(); ++k) e[k] = T[(size_t)ent[k].ar * 20 + ent[k].at];
    for (size_t k = 0; k < rows.size(); ++k) {
      stat[3 * k] = e[rows[k].off];
      stats(&e[rows[k].off + 1], m, &stat[3 * k + 1]);
    }
  } else if (!rows.empty()) {
    // every entry's energy, and each row's statistics, on the device
    DevBuf b_rows, b_ent, b_geo, b_draws, b_out, b_stat;
    if (geo.empty()) geo.push_back(TfGeo{0, 0, 0});
    if (!(s = b_rows.put(rows, st)).ok() || !(s = b_ent.put(ent, st)).ok() || !(s = b_geo.put(geo, st)).ok() ||
        !(s = b_draws.put(draws, st)).ok())
      return s;
    TF_HIPCHK(hipMalloc(&b_out.p, std::max<long long>(off, 1) * sizeof(double)));
    TF_HIPCHK(hipMalloc(&b_stat.p, stat.size() * sizeof(double)));
    hipLaunchKernelGGL(k_tf_rows, dim3(rows.size()), dim3(256), 0, st, D.params, D.rinfo, D.x, n, d_rho,
                       (const TfRow *)b_rows.p, (const TfEnt *)b_ent.p, (const TfGeo *)b_geo.p,
                       (const int *)b_draws.p, (double *)b_out.p, (double *)b_stat.p);
    TF_HIPCHK(hipGetLastError());
    TF_HIPCHK(hipMemcpyAsync(stat.data(), b_stat.p, stat.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    TF_HIPCHK(hipStreamSynchronize(st));
  }
  if (conf_row >= 0) {
    F.stats[0] = stat[3 * conf_row + 1];
    F.stats[1] = stat[3 * conf_row + 2];
    F.done = true;
  }
  table.clear();
  for (size_t k = 0; k < rows.size(); ++k) {
    const TfRow &R = rows[k];
    if ((int)k == conf_row) continue;
    double nat = stat[3 * k], sd[2];
    if (R.kind == 0) { sd[0] = F.stats[0]; sd[1] = F.stats[1]; }
    else { sd[0] = stat[3 * k + 1]; sd[1] = stat[3 * k + 2]; }
    double row[11] = {(double)(R.a + 1), R.kind == 2 ? 0.0 : (double)(R.b + 1), (double)(chain(R.a) + 1),
                      R.kind == 2 ? 0.0 : (double)(chain(R.b) + 1), R.kind == 2 ? 0.0 : R.r, R.ra,
                      R.kind == 2 ? 0.0 : R.rb, nat, sd[0], sd[1], (sd[0] - nat) / sd[1]};
    table.insert(table.end(), row, row + 11);
  }
  return AwStatus();
}

// secondary-structure letters of the n-mer at st from the ssweight helix
// and strand columns (get_nmer_secondary_structure, fb.cpp:5931-5945)
static void nm_ss(const AwSystem &S, int st, char *ss)
{
  for (int o = 0; o < S.P.nf_size; ++o) {
    bool e = S.aps[(size_t)(st + o) * 12 + 4] == 1.0, h = S.aps[(size_t)(st + o) * 12 + 3] == 1.0;
    ss[o] = (e && h) ? '!' : (h ? 'H' : (e ? 'E' : '-'));
  }
  ss[S.P.nf_size] = 0;
}

// one call of the n-mer frustratometer: rows of 7 doubles (i, j, contacts,
// native, decoy mean, decoy std, index; single n-mers i, 0, 0, ...) and, with
// the trap flag, trap records of 7 doubles (row, first n-mer, second n-mer,
// k, backward, threshold, energy; 1-based starts) in the reference's order
static AwStatus nmer_frust(awsem_ctx *c, const double *xres, std::vector<double> &table, std::vector<double> &traps)
{
  const AwSystem &S = *aw_ctx_system(c);
  const FFParams &P = S.P;
  if (!P.nf_flag) return AwStatus::err(AWSEM_E_STATE, "no [Nmer_Frustratometer] section in the coeff file");
  if (P.nf_ndecoys < 1) return AwStatus::err(AWSEM_E_FORMAT, "Nmer_Frustratometer: no decoys");
  if (P.nf_size < 1 || P.nf_size > NM_MAX)
    return AwStatus::err(AWSEM_E_FORMAT, "Nmer_Frustratometer: n-mer size outside 1.." + std::to_string(NM_MAX));
  AwCtxDev D;
  AwStatus st = aw_ctx_device(c, xres, D);
  if (!st.ok()) return st;
  hipStream_t stream = (hipStream_t)D.stream;
  const int n = S.n, s = P.nf_size, m = P.nf_ndecoys;
  AwFrustState &F = aw_ctx_frust(c);
  DevBuf rho_buf;
  TF_HIPCHK(hipMalloc(&rho_buf.p, (size_t)n * sizeof(double)));
  double *d_rho = (double *)rho_buf.p;
  hipLaunchKernelGGL(k_tf_rho, dim3((n + 3) / 4), dim3(256), 0, stream, D.params, D.rinfo, D.x, n, d_rho);
  TF_HIPCHK(hipGetLastError());
  std::vector<double> e;
  table.clear();
  traps.clear();
  auto stats_of = [&](const double *v, double *sd) { stats(v, m, sd); };
  if (P.nf_mode == 1) {
    // compute_singlenmer_frust, fb.cpp:5659-5706
    const int nr = n - s + 1;
    if (nr <= 0) return AwStatus();
    std::vector<int> src((size_t)nr * (m + 1));
    for (int i = 0; i < nr; ++i) {
      src[(size_t)i * (m + 1)] = i;
      for (int d = 0; d < m; ++d) {
        int jr = F.next() % n;
        while (jr + s > n) jr = F.next() % n;
        src[(size_t)i * (m + 1) + 1 + d] = jr;
      }
    }
    DevBuf e1, b_src, b_out;
    TF_HIPCHK(hipMalloc(&e1.p, (size_t)n * 20 * sizeof(double)));
    hipLaunchKernelGGL(k_nm_single, dim3((n * 20 + 255) / 256), dim3(256), 0, stream, D.params, D.rinfo, D.x, n,
                       d_rho, P.nf_cutoff, 1, (double *)e1.p);
    TF_HIPCHK(hipGetLastError());
    if (!(st = b_src.put(src, stream)).ok()) return st;
    TF_HIPCHK(hipMalloc(&b_out.p, src.size() * sizeof(double)));
    long long ne = (long long)src.size();
    hipLaunchKernelGGL(k_nm_single_sum, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, stream, D.rinfo,
                       (const double *)e1.p, s, m, ne, (const int *)b_src.p, (double *)b_out.p);
    TF_HIPCHK(hipGetLastError());
    e.resize(src.size());
    TF_HIPCHK(hipMemcpyAsync(e.data(), b_out.p, e.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    TF_HIPCHK(hipStreamSynchronize(stream));
    for (int i = 0; i < nr; ++i) {
      const double *v = &e[(size_t)i * (m + 1)];
      double sd[2];
      stats_of(v + 1, sd);
      double row[7] = {(double)(i + 1), 0.0, 0.0, v[0], sd[0], sd[1], (sd[0] - v[0]) / sd[1]};
      table.insert(table.end(), row, row + 7);
    }
    return AwStatus();
  }
  // compute_nmer_frust, fb.cpp:5600-5657
  const int ns = n - s;
  if (ns <= 0) return AwStatus();
  // contact counts a block of start rows at a time (at most 2^24 pairs;
  // AWSEM_NMER_BLOCK sets another bound)
  long long maxp = 1 << 24;
  if (const char *e = std::getenv("AWSEM_NMER_BLOCK")) maxp = std::max(1LL, std::atoll(e));
  const int rb = (int)std::max(1LL, std::min((long long)ns, maxp / ns));
  DevBuf b_cnt;
  TF_HIPCHK(hipMalloc(&b_cnt.p, (size_t)rb * ns * sizeof(int)));
  std::vector<int> cnt((size_t)rb * ns);
  const int nt = P.nf_trap ? 4 : 0;  // trap entries: (i, j) and (j, i), forward and backward
  std::vector<NmRow> rows;
  std::vector<int> draws;  // the decoys' n-mer starts, two per decoy
  long long off = 0;
  for (int i = 0; i < ns; ++i) {
    if (i % rb == 0) {
      const int i1 = std::min(ns, i + rb);
      const long long np = (long long)(i1 - i) * ns;
      hipLaunchKernelGGL(k_nm_contacts, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, stream, D.params, D.rinfo,
                         D.x, n, s, i, i1, (int *)b_cnt.p);
      TF_HIPCHK(hipGetLastError());
      TF_HIPCHK(hipMemcpyAsync(cnt.data(), b_cnt.p, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, stream));
      TF_HIPCHK(hipStreamSynchronize(stream));
    }
    for (int j = i + 1; j < ns; ++j) {
      int k = cnt[(size_t)(i % rb) * ns + j];
      if (!(k >= P.nf_contacts && j - i >= s)) continue;
      rows.push_back(NmRow{i, j, 1 + m + nt, k, off, (long long)draws.size(), m, nt});
      for (int d = 0; d < m; ++d) {
        int ir = F.next() % n, jr = F.next() % n;
        if (ir > jr) std::swap(ir, jr);
        while (ir + s > n || jr + s > n || jr - ir < s) {
          ir = F.next() % n;
          jr = F.next() % n;
          if (ir > jr) std::swap(ir, jr);
        }
        draws.push_back(ir);
        draws.push_back(jr);
      }
      off += 1 + m + nt;
    }
  }
  if (rows.empty()) return AwStatus();
  // every entry's energy, and per row the native, decoy mean and std and
  // the four trap energies, on the device
  DevBuf b_rows, b_draws, b_out, b_stat;
  if (!(st = b_rows.put(rows, stream)).ok() || !(st = b_draws.put(draws, stream)).ok()) return st;
  TF_HIPCHK(hipMalloc(&b_out.p, (size_t)off * sizeof(double)));
  TF_HIPCHK(hipMalloc(&b_stat.p, rows.size() * 7 * sizeof(double)));
  hipLaunchKernelGGL(k_nm_pairs, dim3((unsigned)rows.size()), dim3(256), 0, stream, D.params, D.rinfo, D.x, n, s,
                     d_rho, (const NmRow *)b_rows.p, (const int *)b_draws.p, (double *)b_out.p, (double *)b_stat.p);
  TF_HIPCHK(hipGetLastError());
  e.resize(rows.size() * 7);
  TF_HIPCHK(hipMemcpyAsync(e.data(), b_stat.p, e.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  TF_HIPCHK(hipStreamSynchronize(stream));
  std::vector<char> ssg(s + 1), ssk(s + 1);
  for (size_t r = 0; r < rows.size(); ++r) {
    const NmRow &R = rows[r];
    const double *v = &e[7 * r];  // native, mean, std, traps (i, j) fw / bw, (j, i) fw / bw
    const double sd[2] = {v[1], v[2]};
    if (nt) {
      // compute_nmer_traps (fb.cpp:5786-5893) for (i, j), then (j, i)
      const double thr = v[0] - P.nf_trap_sigma * sd[1];
      for (int dir = 0; dir < 2; ++dir) {
        const int f = dir ? R.b : R.a, g = dir ? R.a : R.b;
        nm_ss(S, g, ssg.data());
        for (int bw = 0; bw < 2; ++bw) {
          const double et = v[3 + 2 * dir + bw];
          for (int k = 0; k < ns; ++k) {
            if (std::abs(k - g) <= s || (std::abs(k - f) <= s && f != k)) continue;
            nm_ss(S, k, ssk.data());
            int sdist = 0;
            for (int o = 0; o < s; ++o) sdist += ssg[o] != ssk[o];
            if (sdist > s * (1 - P.nf_ss_frac)) continue;
            if (et < thr) {
              double q[7] = {(double)r, (double)(f + 1), (double)(g + 1), (double)(k + 1), (double)bw, thr, et};
              traps.insert(traps.end(), q, q + 7);
            }
          }
        }
      }
    }
    double row[7] = {(double)(R.a + 1), (double)(R.b + 1), (double)R.ncont, v[0], sd[0], sd[1], (sd[0] - v[0]) / sd[1]};
    table.insert(table.end(), row, row + 7);
  }
  return AwStatus();
}

// the (memory, relative pair) items of a library (the pairs of
// compute_decoy_memory_potential, fb.cpp:4478-4496: residue di from 0 to
// len - minSep - 1, partner dj from di + minSep to len - 1, at most
// di + maxSep); memory distances as Fragment_Memory::Rf (fm.cpp:146-163)
static void ff_items(const AwSystem &S, const std::vector<AwMemory> &lib, std::vector<FfItem> &out)
{
  const int minsep = S.fmg.min_sep(), maxsep = S.fmg.max_sep();
  auto R = [](const double *a, const double *b) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
  };
  out.clear();
  for (int k = 0; k < (int)lib.size(); ++k) {
    const AwMemory &m = lib[k];
    for (int di = 0; di < m.len - minsep; ++di) {
      int je = m.len - 1;
      if (maxsep != -1) je = std::min(je, di + maxsep);
      for (int dj = di + minsep; dj <= je; ++dj) {
        FfItem I;
        I.k = k;
        I.di = di;
        I.dj = dj;
        I.gly = (m.se[di] == 'G' ? 1 : 0) | (m.se[dj] == 'G' ? 2 : 0);
        const double *ai = &m.ca[3 * di], *aj = &m.ca[3 * dj], *bi = &m.cb[3 * di], *bj = &m.cb[3 * dj];
        I.rf[0] = R(ai, aj);
        I.rf[1] = R(ai, bj);
        I.rf[2] = R(aj, bi);
        I.rf[3] = R(bj, bi);
        out.push_back(I);
      }
    }
  }
}

// one fragment-frustratometer calculation (fb.cpp:7537-7576): rows n x 3
// (index, native - mean, std per residue); in read mode, at the call that
// computes the decoy statistics, gen = n natives then n x n_decoys decoy
// energies (residue-major) and *gen_written = 1
static AwStatus frag_frust(awsem_ctx *c, const double *xres, std::vector<double> &rows, std::vector<double> *gen,
                           int *gen_written)
{
  const AwSystem &S = *aw_ctx_system(c);
  const FFParams &P = S.P;
  *gen_written = 0;
  if (!P.ff_flag) return AwStatus::err(AWSEM_E_STATE, "no [Fragment_Frustratometer] section in the coeff file");
  if (P.ff_ncalc - 1 > 65535 || (int)S.ff_mems.size() > 65535)
    return AwStatus::err(AWSEM_E_FORMAT, "Fragment_Frustratometer: more than 65535 decoy calculations or structures");
  AwCtxDev D;
  AwStatus st = aw_ctx_device(c, xres, D);
  if (!st.ok()) return st;
  hipStream_t stream = (hipStream_t)D.stream;
  const int n = S.n, nc = P.ff_ncalc, nfm = (int)S.mems.size(), nd = (int)S.ff_mems.size();
  AwFrustState &F = aw_ctx_frust(c);
  if (F.ff_pos.empty())
    for (const AwMemory &m : S.ff_mems) F.ff_pos.push_back(m.pos);
  if (!F.ff_ready) {
    // gamma by separation and types (Gamma_Array::getGamma, fm.cpp:400-424)
    int dmax = 0;
    for (const AwMemory &m : S.mems) dmax = std::max(dmax, m.len);
    for (const AwMemory &m : S.ff_mems) dmax = std::max(dmax, m.len);
    std::vector<double> gsep((size_t)(dmax + 1) * 400);
    for (int d = 0; d <= dmax; ++d)
      for (int a = 0; a < 20; ++a)
        for (int b = 0; b < 20; ++b) gsep[(size_t)d * 400 + a * 20 + b] = S.fmg.get(a, b, 0, d);
    std::vector<FfItem> nat_items, dec_items;
    ff_items(S, S.mems, nat_items);
    ff_items(S, S.ff_mems, dec_items);
    std::vector<int> nat_pos(std::max(nfm, 1));
    std::vector<double> nat_w(std::max(nfm, 1)), dec_w(std::max(nd, 1));
    for (int k = 0; k < nfm; ++k) { nat_pos[k] = S.mems[k].pos; nat_w[k] = S.mems[k].weight; }
    for (int k = 0; k < nd; ++k) dec_w[k] = S.ff_mems[k].weight;
    if (!(st = put_static(gsep, F.ff_dev[0])).ok() || !(st = put_static(nat_items, F.ff_dev[1])).ok() ||
        !(st = put_static(nat_pos, F.ff_dev[2])).ok() || !(st = put_static(nat_w, F.ff_dev[3])).ok() ||
        !(st = put_static(dec_items, F.ff_dev[4])).ok() || !(st = put_static(dec_w, F.ff_dev[5])).ok())
      return st;
    F.ff_n_nat = (int)nat_items.size();
    F.ff_n_dec = (int)dec_items.size();
    F.ff_ready = true;
  }
  DevBuf b_err, b_E;
  std::vector<int> zero(1, 0);
  if (!(st = b_err.put(zero, stream)).ok()) return st;
  const int ncalc_dev = P.ff_mode == 0 ? nc : 1;  // read mode: the natives only
  TF_HIPCHK(hipMalloc(&b_E.p, (size_t)ncalc_dev * n * sizeof(double)));
  TF_HIPCHK(hipMemsetAsync(b_E.p, 0, (size_t)ncalc_dev * n * sizeof(double), stream));
  double *E = (double *)b_E.p;
  int *d_err = (int *)b_err.p;
  // a library's energies for `calcs` calculations (LDS accumulation when
  // the residues fit)
  auto launch_mem = [&](const void *items, int n_items, const void *pos, const void *w, int nm, int calcs,
                        double *out) {
    if (n <= FF_LDS_MAX)
      hipLaunchKernelGGL(k_ff_mem_lds, dim3((unsigned)((n_items + FF_CHUNK - 1) / FF_CHUNK), (unsigned)calcs),
                         dim3(256), (size_t)n * sizeof(double), stream, D.params, D.rinfo, D.x, n,
                         (const FfItem *)items, n_items, (const int *)pos, (const double *)w, nm,
                         (const double *)F.ff_dev[0], out, d_err);
    else
      hipLaunchKernelGGL(k_ff_mem, dim3((unsigned)((n_items + 255) / 256), (unsigned)calcs), dim3(256), 0, stream,
                         D.params, D.rinfo, D.x, n, (const FfItem *)items, n_items, (const int *)pos,
                         (const double *)w, nm, (const double *)F.ff_dev[0], out, d_err);
    return hipGetLastError();
  };
  if (F.ff_n_nat > 0) TF_HIPCHK(launch_mem(F.ff_dev[1], F.ff_n_nat, F.ff_dev[2], F.ff_dev[3], nfm, 1, E));
  DevBuf b_dp, b_dx, b_dinfo, b_G;
  if (P.ff_mode == 0) {
    // randomize_decoys after every calculation (fb.cpp:4565-4611, 7546-7553):
    // calculation c >= 1 sees the placement drawn after calculation c - 1
    std::vector<int> pos((size_t)(nc - 1) * std::max(nd, 1));
    for (int r = 0; r < nc; ++r)
      for (int k = 0; k < nd; ++k) {
        int p = F.next() % (n - S.ff_mems[k].len + 1);
        F.ff_pos[k] = p;
        if (r + 1 < nc) pos[(size_t)r * nd + k] = p;
      }
    if (F.ff_n_dec > 0) {
      if (!(st = b_dp.put(pos, stream)).ok()) return st;
      TF_HIPCHK(launch_mem(F.ff_dev[4], F.ff_n_dec, b_dp.p, F.ff_dev[5], nd, nc - 1, E + n));
    }
  }
  const bool gen_now = P.ff_mode == 1 && !F.ff_gen_done;
  if (gen_now && F.ff_n_nat > 0) {
    // the decoy structures' coordinates: CA then CB per decoy
    std::vector<double> dxyz;
    std::vector<int> dinfo;
    for (const AwMemory &m : S.ff_mems) {
      dinfo.push_back(m.pos);
      dinfo.push_back(m.len);
      dinfo.push_back((int)dxyz.size());
      dxyz.insert(dxyz.end(), m.ca.begin(), m.ca.end());
      dxyz.insert(dxyz.end(), m.cb.begin(), m.cb.end());
    }
    if (!(st = b_dx.put(dxyz, stream)).ok() || !(st = b_dinfo.put(dinfo, stream)).ok()) return st;
    TF_HIPCHK(hipMalloc(&b_G.p, (size_t)nd * n * sizeof(double)));
    TF_HIPCHK(hipMemsetAsync(b_G.p, 0, (size_t)nd * n * sizeof(double), stream));
    hipLaunchKernelGGL(k_ff_gen, dim3((unsigned)((F.ff_n_nat + 255) / 256), (unsigned)nd), dim3(256), 0, stream,
                       D.params, D.rinfo, n, (const FfItem *)F.ff_dev[1], F.ff_n_nat, (const int *)F.ff_dev[2],
                       (const double *)F.ff_dev[3], (const double *)F.ff_dev[0], (const double *)b_dx.p,
                       (const int *)b_dinfo.p,
                       (double *)b_G.p, d_err);
    TF_HIPCHK(hipGetLastError());
  }
  std::vector<double> e((size_t)ncalc_dev * n), g(gen_now ? (size_t)nd * n : 0);
  int herr = 0;
  TF_HIPCHK(hipMemcpyAsync(e.data(), E, e.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  if (gen_now && !g.empty())
    TF_HIPCHK(hipMemcpyAsync(g.data(), b_G.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  TF_HIPCHK(hipMemcpyAsync(&herr, d_err, sizeof(int), hipMemcpyDeviceToHost, stream));
  TF_HIPCHK(hipStreamSynchronize(stream));
  if (herr) return AwStatus::err(AWSEM_E_FORMAT, "Decoy Memory: Interaction between residues of different chains");
  rows.assign((size_t)n * 3, 0.0);
  if (P.ff_mode == 0) {
    // compute_fragment_frustration, fb.cpp:4613-4689
    for (int i = 0; i < n; ++i) {
      e[i] /= nfm;
      for (int d = 1; d < nc; ++d) e[(size_t)d * n + i] /= nd;
    }
    for (int i = 0; i < n; ++i) {
      double mean = 0.0, var = 0.0;
      for (int d = 1; d < nc; ++d) mean += e[(size_t)d * n + i];
      mean /= (nc - 1);
      for (int d = 1; d < nc; ++d) var += std::pow(e[(size_t)d * n + i] - mean, 2);
      var /= (nc - 1);
      rows[3 * i] = (e[i] - mean) / std::sqrt(var);
      rows[3 * i + 1] = e[i] - mean;
      rows[3 * i + 2] = std::sqrt(var);
    }
    return AwStatus();
  }
  if (gen_now) {
    // the per-residue decoy statistics (fb.cpp:4804-4830), once: at step 0
    // in the reference, at the first calculation here
    F.ff_mean.assign(n, 0.0);
    F.ff_var.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
      double mean = 0.0, var = 0.0;
      for (int d = 0; d < nd; ++d) mean += g[(size_t)d * n + i];
      mean /= nd;
      for (int d = 0; d < nd; ++d) var += std::pow(g[(size_t)d * n + i] - mean, 2);
      var /= nd;
      F.ff_mean[i] = mean;
      F.ff_var[i] = var;
    }
    F.ff_gen_done = true;
    if (gen) {
      gen->assign((size_t)n * (nd + 1), 0.0);
      for (int i = 0; i < n; ++i) (*gen)[i] = e[i];
      for (int i = 0; i < n; ++i)
        for (int d = 0; d < nd; ++d) (*gen)[n + (size_t)i * nd + d] = g[(size_t)d * n + i];
      *gen_written = 1;
    }
  }
  for (int i = 0; i < n; ++i) {
    rows[3 * i] = (e[i] - F.ff_mean[i]) / std::sqrt(F.ff_var[i]);
    rows[3 * i + 1] = e[i] - F.ff_mean[i];
    rows[3 * i + 2] = std::sqrt(F.ff_var[i]);
  }
  return AwStatus();
}

// the generator's state, kept so that a call whose output does not fit
// leaves the stream where it was
struct TfSaved {
  uint32_t rng[34];
  int k;
  bool done;
  double stats[2];
  explicit TfSaved(const AwFrustState &F) : k(F.rng_k), done(F.done)
  {
    std::memcpy(rng, F.rng, sizeof rng);
    stats[0] = F.stats[0];
    stats[1] = F.stats[1];
  }
  void restore(AwFrustState &F) const
  {
    std::memcpy(F.rng, rng, sizeof rng);
    F.rng_k = k;
    F.done = done;
    F.stats[0] = stats[0];
    F.stats[1] = stats[1];
  }
};

extern "C" {

int awsem_tert_frust(awsem_ctx *c, const double *xres, double *rows, int cap, int *n_rows)
{
  if (!c || !xres || !n_rows || cap < 0 || (cap > 0 && !rows)) return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(c);
  if (!S || S->n <= 0) return AWSEM_E_ARG;
  for (size_t k = 0; k < (size_t)S->n * 9; ++k)
    if (!std::isfinite(xres[k])) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates"));
  std::vector<double> t;
  const TfSaved saved(aw_ctx_frust(c));
  AwStatus s = tert_frust(c, xres, t);
  if (!s.ok()) return aw_ctx_fail(c, s);
  int nr = (int)(t.size() / 11);
  *n_rows = nr;
  if (nr > cap) {
    saved.restore(aw_ctx_frust(c));
    return aw_ctx_fail(c, AwStatus::err(AWSEM_E_ARG, "awsem_tert_frust: more rows than cap"));
  }
  if (nr) std::memcpy(rows, t.data(), t.size() * sizeof(double));
  return AWSEM_OK;
}

// the reference's output at step % tert_frust_output_freq == 0
// (fb.cpp:7586-7596, 5057-5189; the files opened and headed at
// construction, fb.cpp:855-867)
int awsem_tert_frust_write(awsem_ctx *c, const double *xres, int64_t step)
{
  if (!c || !xres) return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(c);
  if (!S || S->n <= 0) return AWSEM_E_ARG;
  const FFParams &P = S->P;
  if (!P.tf_flag) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_STATE, "no [Tertiary_Frustratometer] section in the coeff file"));
  if (P.tf_freq <= 0 || step % P.tf_freq != 0) return AWSEM_OK;
  AwFrustState &F = aw_ctx_frust(c);
  const std::string &dir = aw_ctx_dir(c);
  if (!F.dat) {
    F.dat = std::fopen((dir + "/tertiary_frustration.dat").c_str(), "w");
    F.tcl = std::fopen((dir + "/tertiary_frustration.tcl").c_str(), "w");
    if (!F.dat || !F.tcl) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_FILE, "cannot open tertiary_frustration.dat/.tcl"));
    if (P.tf_mode != 2)
      std::fprintf(F.dat, "# i j i_chain j_chain xi yi zi xj yj zj r_ij rho_i rho_j a_i a_j native_energy "
                          "<decoy_energies> std(decoy_energies) f_ij\n");
    else
      std::fprintf(F.dat, "# i i_chain xi yi zi rho_i a_i native_energy <decoy_energies> std(decoy_energies) f_i\n");
  }
  std::vector<double> t;
  for (size_t k = 0; k < (size_t)S->n * 9; ++k)
    if (!std::isfinite(xres[k])) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates"));
  AwStatus s = tert_frust(c, xres, t);
  if (!s.ok()) return aw_ctx_fail(c, s);
  std::fprintf(F.dat, "# timestep: %d\n", (int)step);
  std::fprintf(F.tcl, "# timestep: %d\n", (int)step);
  int n = S->n, atomselect = 0;
  for (size_t k = 0; k + 11 <= t.size(); k += 11) {
    const double *q = &t[k];
    int i = (int)q[0] - 1;
    const double *xi = site(*S, xres, i);
    if (P.tf_mode != 2) {
      int j = (int)q[1] - 1;
      const double *xj = site(*S, xres, j);
      std::fprintf(F.dat, "%5d %5d %3d %3d %8.3f %8.3f %8.3f %8.3f %8.3f %8.3f %8.3f %8.3f %8.3f %c %c %8.3f %8.3f %8.3f %8.3f\n",
                   i + 1, j + 1, (int)q[2], (int)q[3], xi[0], xi[1], xi[2], xj[0], xj[1], xj[2], q[4], q[5], q[6],
                   S->se[i], S->se[j], q[7], q[8], q[9], q[10]);
      if (q[10] > 0.78 || q[10] < -1) {
        std::fprintf(F.tcl, "set sel%d [atomselect top \"resid %d and name CA\"]\n", i, i + 1);
        std::fprintf(F.tcl, "set sel%d [atomselect top \"resid %d and name CA\"]\n", j, j + 1);
        std::fprintf(F.tcl, "lassign [atomselect%d get {x y z}] pos1\n", atomselect++);
        std::fprintf(F.tcl, "lassign [atomselect%d get {x y z}] pos2\n", atomselect++);
        std::fprintf(F.tcl, q[10] > 0.78 ? "draw color green\n" : "draw color red\n");
        if (q[4] < P.well_r_max[0]) std::fprintf(F.tcl, "draw line $pos1 $pos2 style solid width 1\n");
        else std::fprintf(F.tcl, "draw line $pos1 $pos2 style dashed width 2\n");
      }
    } else {
      std::fprintf(F.dat, "%5d %5d %8.3f %8.3f %8.3f %8.3f %c %8.3f %8.3f %8.3f %8.3f\n", i + 1, (int)q[2], xi[0], xi[1],
                   xi[2], q[5], S->se[i], q[7], q[8], q[9], q[10]);
      atomselect += 1;
      std::fprintf(F.tcl, "mol addrep 0\n");
      std::fprintf(F.tcl, "mol modselect %d 0 resid %d\n", atomselect, i + 1);
      std::fprintf(F.tcl, "mol modstyle %d 0 VDW %f 12.000000\n", atomselect, 0.5 * std::fabs(q[10]));
      std::fprintf(F.tcl, "mol modmaterial %d 0 Transparent\n", atomselect);
      std::fprintf(F.tcl, q[10] > 0.0 ? "mol modcolor %d 0 ColorID 7\n" : "mol modcolor %d 0 ColorID 1\n", atomselect);
    }
  }
  std::fprintf(F.tcl, "mol modselect 0 top \"all\"\n");
  std::fprintf(F.tcl, "mol modstyle 0 top newcartoon\n");
  std::fprintf(F.tcl, "mol modcolor 0 top colorid 15\n");
  std::fflush(F.dat);
  std::fflush(F.tcl);
  (void)n;
  return AWSEM_OK;
}

int awsem_tert_frust_every(const awsem_ctx *c, int64_t *every)
{
  if (!c || !every) return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(c);
  if (!S) return AWSEM_E_ARG;
  *every = S->P.tf_flag ? (int64_t)S->P.tf_freq : 0;
  return AWSEM_OK;
}

int awsem_nmer_frust(awsem_ctx *c, const double *xres, double *rows, int cap, int *n_rows, double *traps, int tcap,
                     int *n_traps)
{
  if (!c || !xres || !n_rows || !n_traps || cap < 0 || tcap < 0 || (cap > 0 && !rows) || (tcap > 0 && !traps))
    return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(c);
  if (!S || S->n <= 0) return AWSEM_E_ARG;
  for (size_t k = 0; k < (size_t)S->n * 9; ++k)
    if (!std::isfinite(xres[k])) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates"));
  std::vector<double> t, tr;
  const TfSaved saved(aw_ctx_frust(c));
  AwStatus s = nmer_frust(c, xres, t, tr);
  if (!s.ok()) return aw_ctx_fail(c, s);
  int nr = (int)(t.size() / 7), nt = (int)(tr.size() / 7);
  *n_rows = nr;
  *n_traps = nt;
  if (nr > cap || nt > tcap) {
    saved.restore(aw_ctx_frust(c));
    return aw_ctx_fail(c, AwStatus::err(AWSEM_E_ARG, "awsem_nmer_frust: more rows or traps than cap"));
  }
  if (nr) std::memcpy(rows, t.data(), t.size() * sizeof(double));
  if (nt) std::memcpy(traps, tr.data(), tr.size() * sizeof(double));
  return AWSEM_OK;
}

// the reference's output at step % nmer_frust_output_freq == 0
// (fb.cpp:7592-7603, 5600-5706, 5855-5890; the files opened and headed at
// construction, fb.cpp:885-896)
int awsem_nmer_frust_write(awsem_ctx *c, const double *xres, int64_t step)
{
  if (!c || !xres) return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(c);
  if (!S || S->n <= 0) return AWSEM_E_ARG;
  const FFParams &P = S->P;
  if (!P.nf_flag) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_STATE, "no [Nmer_Frustratometer] section in the coeff file"));
  if (P.nf_freq <= 0 || step % P.nf_freq != 0) return AWSEM_OK;
  AwFrustState &F = aw_ctx_frust(c);
  const std::string &dir = aw_ctx_dir(c);
  if (!F.nm_dat) {
    F.nm_dat = std::fopen((dir + "/nmer_frustration.dat").c_str(), "w");
    F.nm_tcl = std::fopen((dir + "/nmer_frustration.tcl").c_str(), "w");
    if (P.nf_trap) F.nm_trap = std::fopen((dir + "/nmer_traps.dat").c_str(), "w");
    if (!F.nm_dat || !F.nm_tcl || (P.nf_trap && !F.nm_trap))
      return aw_ctx_fail(c, AwStatus::err(AWSEM_E_FILE, "cannot open nmer_frustration.dat/.tcl or nmer_traps.dat"));
    if (P.nf_mode == 0)
      std::fprintf(F.nm_dat, "# i j ncontacts a_i a_j native_energy <decoy_energies> std(decoy_energies) f_ij\n");
    else
      std::fprintf(F.nm_dat, "# i a_i native_energy <decoy_energies> std(decoy_energies) f_ij\n");
    if (P.nf_trap)
      std::fprintf(F.nm_trap, "# i a_i ss_i j a_j ss_j threshold_energy k a_k ss_k direction trap_energy\n");
  }
  for (size_t k = 0; k < (size_t)S->n * 9; ++k)
    if (!std::isfinite(xres[k])) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates"));
  std::vector<double> t, tr;
  AwStatus s = nmer_frust(c, xres, t, tr);
  if (!s.ok()) return aw_ctx_fail(c, s);
  const int ns = P.nf_size;
  auto seq = [&](int st, int bw) {
    std::string q(ns, ' ');
    for (int o = 0; o < ns; ++o) q[o] = S->se[bw ? st + ns - o : st + o];
    return q;
  };
  auto ss = [&](int st) {
    std::vector<char> b(ns + 1);
    nm_ss(*S, st, b.data());
    return std::string(b.data());
  };
  FILE *dat = F.nm_dat, *tcl = F.nm_tcl;
  std::fprintf(dat, "# timestep: %d\n", (int)step);
  std::fprintf(tcl, "# timestep: %d\n", (int)step);
  if (P.nf_mode == 1) {
    std::fprintf(tcl, "mol modselect 0 top \"all\"\n");
    std::fprintf(tcl, "mol modstyle 0 top newcartoon\n");
    std::fprintf(tcl, "mol modcolor 0 top colorid 15\n");
    int atomselect = 0;
    for (size_t k = 0; k + 7 <= t.size(); k += 7) {
      const double *q = &t[k];
      int i = (int)q[0] - 1;
      std::fprintf(dat, "%d %s %f %f %f %f\n", i + 1, seq(i, 0).c_str(), q[3], q[4], q[5], q[6]);
      if (q[6] > P.nf_min_thr || q[6] < P.nf_high_thr) {
        atomselect += 1;
        std::fprintf(tcl, "mol addrep 0\n");
        std::fprintf(tcl, "mol modselect %d 0 resid %d to %d\n", atomselect, i + 1, i + ns);
        std::fprintf(tcl, "mol modstyle %d 0 VDW %f 12.000000\n", atomselect, 0.5 * std::fabs(q[6]));
        std::fprintf(tcl, "mol modmaterial %d 0 Transparent\n", atomselect);
        if (q[6] > P.nf_min_thr) std::fprintf(tcl, "mol modcolor %d 0 ColorID 7\n", atomselect);
        else std::fprintf(tcl, "mol modcolor %d 0 ColorID 1\n", atomselect);
      }
    }
  } else {
    int atomselect = 0;
    size_t tk = 0;
    for (size_t k = 0; k + 7 <= t.size(); k += 7) {
      const double *q = &t[k];
      const int row = (int)(k / 7), i = (int)q[0] - 1, j = (int)q[1] - 1;
      for (; tk + 7 <= tr.size() && (int)tr[tk] == row; tk += 7) {
        const double *r = &tr[tk];
        int f = (int)r[1] - 1, g = (int)r[2] - 1, kk = (int)r[3] - 1, bw = (int)r[4];
        std::fprintf(F.nm_trap, bw ? "%d %s %s %d %s %s %f %d %s <-- %s %f \n" : "%d %s %s %d %s %s %f %d %s --> %s %f \n",
                     f + 1, seq(f, 0).c_str(), ss(f).c_str(), g + 1, seq(g, 0).c_str(), ss(g).c_str(), r[5], kk + 1,
                     seq(kk, bw).c_str(), ss(kk).c_str(), r[6]);
        if (P.nf_draw_trap) {
          if (f == kk) {
            std::fprintf(tcl, "mol addrep 0\n");
            std::fprintf(tcl, "mol modselect %d 0 resid %d to %d\n", F.nm_rep, f + 1, f + ns);
            std::fprintf(tcl, "mol modcolor %d 0 ColorID 11\n", F.nm_rep);
            std::fprintf(tcl, "mol modstyle %d 0 NewCartoon 0.350000 10.000000 4.100000 0\n", F.nm_rep);
            F.nm_rep++;
          } else {
            std::fprintf(tcl, "set sel%d [atomselect top \"resid %d and name CA\"]\n", f + ns / 2, f + 1 + ns / 2);
            std::fprintf(tcl, "set sel%d [atomselect top \"resid %d and name CA\"]\n", kk + ns / 2, kk + 1 + ns / 2);
            std::fprintf(tcl, "lassign [atomselect%d get {x y z}] pos1\n", atomselect++);
            std::fprintf(tcl, "lassign [atomselect%d get {x y z}] pos2\n", atomselect++);
            std::fprintf(tcl, "draw color purple\n");
            std::fprintf(tcl, bw ? "draw line $pos1 $pos2 style dashed width 1\n" : "draw line $pos1 $pos2 style solid width 1\n");
          }
        }
      }
      std::fprintf(dat, "%d %d %d %s %s %f %f %f %f\n", i + 1, j + 1, (int)q[2], seq(i, 0).c_str(), seq(j, 0).c_str(),
                   q[3], q[4], q[5], q[6]);
      if (q[6] > P.nf_min_thr || q[6] < P.nf_high_thr || P.nf_neutral) {
        std::fprintf(tcl, "set sel%d [atomselect top \"resid %d and name CA\"]\n", i + ns / 2, i + 1 + ns / 2);
        std::fprintf(tcl, "set sel%d [atomselect top \"resid %d and name CA\"]\n", j + ns / 2, j + 1 + ns / 2);
        std::fprintf(tcl, "lassign [atomselect%d get {x y z}] pos1\n", atomselect++);
        std::fprintf(tcl, "lassign [atomselect%d get {x y z}] pos2\n", atomselect++);
        if (q[6] > P.nf_min_thr) std::fprintf(tcl, "draw color green\n");
        else if (q[6] < P.nf_high_thr) std::fprintf(tcl, "draw color red\n");
        else std::fprintf(tcl, "draw color blue\n");
        std::fprintf(tcl, "draw line $pos1 $pos2 style solid width 1\n");
      }
    }
    std::fprintf(tcl, "mol modselect 0 top \"all\"\n");
    std::fprintf(tcl, "mol modstyle 0 top newcartoon\n");
    std::fprintf(tcl, "mol modcolor 0 top colorid 15\n");
  }
  std::fflush(dat);
  std::fflush(tcl);
  if (F.nm_trap) std::fflush(F.nm_trap);
  return AWSEM_OK;
}

int awsem_nmer_frust_every(const awsem_ctx *c, int64_t *every)
{
  if (!c || !every) return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(c);
  if (!S) return AWSEM_E_ARG;
  *every = S->P.nf_flag ? (int64_t)S->P.nf_freq : 0;
  return AWSEM_OK;
}

int awsem_frag_frust_info(const awsem_ctx *c, int *mode, int *n_calcs, int *n_decoys, int64_t *every)
{
  if (!c || !mode || !n_calcs || !n_decoys || !every) return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(c);
  if (!S) return AWSEM_E_ARG;
  *mode = S->P.ff_flag ? S->P.ff_mode : -1;
  *n_calcs = S->P.ff_flag ? S->P.ff_ncalc : 0;
  *n_decoys = (int)S->ff_mems.size();
  *every = S->P.ff_flag ? (int64_t)S->P.ff_freq : 0;
  return AWSEM_OK;
}

int awsem_frag_frust(awsem_ctx *c, const double *xres, double *rows, double *gen, int *gen_written)
{
  if (!c || !xres || !rows || !gen_written) return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(c);
  if (!S || S->n <= 0) return AWSEM_E_ARG;
  for (size_t k = 0; k < (size_t)S->n * 9; ++k)
    if (!std::isfinite(xres[k])) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates"));
  std::vector<double> r, g;
  AwStatus s = frag_frust(c, xres, r, gen ? &g : nullptr, gen_written);
  if (!s.ok()) return aw_ctx_fail(c, s);
  std::memcpy(rows, r.data(), r.size() * sizeof(double));
  if (*gen_written) std::memcpy(gen, g.data(), g.size() * sizeof(double));
  return AWSEM_OK;
}

// the reference's output at step % frag_frust_output_freq == 0
// (fb.cpp:7537-7576, 4676-4688, 4786-4802; the five files opened at
// construction, fb.cpp:794-800)
int awsem_frag_frust_write(awsem_ctx *c, const double *xres, int64_t step)
{
  if (!c || !xres) return AWSEM_E_ARG;
  const AwSystem *S = aw_ctx_system(c);
  if (!S || S->n <= 0) return AWSEM_E_ARG;
  const FFParams &P = S->P;
  if (!P.ff_flag) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_STATE, "no [Fragment_Frustratometer] section in the coeff file"));
  if (P.ff_freq <= 0 || step % P.ff_freq != 0) return AWSEM_OK;
  AwFrustState &F = aw_ctx_frust(c);
  const std::string &dir = aw_ctx_dir(c);
  static const char *names[5] = {"fragment_frustration.dat", "fragment_frustration_gap.dat",
                                 "fragment_frustration_variance.dat", "fragment_frustration_decoy.dat",
                                 "fragment_frustration_native.dat"};
  if (!F.ff_out[0])
    for (int k = 0; k < 5; ++k) {
      F.ff_out[k] = std::fopen((dir + "/" + names[k]).c_str(), "w");
      if (!F.ff_out[k]) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_FILE, std::string("cannot open ") + names[k]));
    }
  for (size_t k = 0; k < (size_t)S->n * 9; ++k)
    if (!std::isfinite(xres[k])) return aw_ctx_fail(c, AwStatus::err(AWSEM_E_NONFINITE, "non-finite coordinates"));
  std::vector<double> r, g;
  int gw = 0;
  AwStatus s = frag_frust(c, xres, r, &g, &gw);
  if (!s.ok()) return aw_ctx_fail(c, s);
  const int n = S->n, nd = (int)S->ff_mems.size();
  if (gw) {
    for (int i = 0; i < n; ++i) std::fprintf(F.ff_out[4], "%f ", g[i]);
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < nd; ++d) std::fprintf(F.ff_out[3], " %f\n", g[n + (size_t)i * nd + d]);
  }
  for (int i = 0; i < n; ++i) {
    std::fprintf(F.ff_out[0], "%f ", r[3 * i]);
    std::fprintf(F.ff_out[1], "%f ", r[3 * i + 1]);
    std::fprintf(F.ff_out[2], "%f ", r[3 * i + 2]);
  }
  for (int k = 0; k < 3; ++k) std::fprintf(F.ff_out[k], "\n");
  for (FILE *f : F.ff_out) std::fflush(f);
  return AWSEM_OK;
}

}  // extern "C"
