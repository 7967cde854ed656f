// awsem_extra.h -- the coeff-toggled FixBackbone terms on the device:
// AMH-Go (fb.cpp:3854-3992), contact restraints (8059-8061, 8137-8177),
// vector fragment memory (3994-4077), the solvent-separated barrier
// (6237-6305), the membrane (6107-6235) and Debye-Hueckel (6307-6363).
// Phosphorylation changes the water and burial gammas inside the pair
// passes (RI_PHOS, FFParams::phosph_gamma).
//
// One pass (xt_pass) after the list build evaluates all of them: per-residue
// membrane terms, the item lists the host built (AMH-Go pairs grouped by
// their owning atom, contact restraints, charged pairs, vector-memory
// pairs), and the barrier over the residue pairs of the Verlet list (or
// every pair).  AMH-Go is many-body only through its owning atom (E_a^p):
// one lane sums E_a over a's items, then adds the forces of the same items
// scaled by a's own factor, so no second pass or global accumulator is
// needed.  The arithmetic is FP64 in both precisions: these terms are off in
// the batched workloads and a few per residue, and FP64 leaves their slopes
// (acos, E_a^(p-1)) unconditioned by rounding.  Forces are added with
// atomics into the replica's force array; energies go to the accumulators
// ered[N_EACC + 0..5].
#pragma once
#include "awsem_terms.h"

// accumulator slots after the twelve of EAcc
enum { XT_GO = 0, XT_VFM, XT_CR, XT_MEMB, XT_SSB, XT_DH, N_XT };

struct XtAcc {
  double e[N_XT];
};

AW_DEV double xt_site_sq(const double *a, const double *b, double d[3])
{
  d[0] = a[0] - b[0]; d[1] = a[1] - b[1]; d[2] = a[2] - b[2];
  return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

template <class T>
AW_DEV void xt_fadd(const View<T> &V, int kind, int i, double s, const double d[3])
{
  double *p = FA(V, kind, i);
  aadd(p + 0, s * d[0]);
  aadd(p + 1, s * d[1]);
  aadd(p + 2, s * d[2]);
}

template <class T>
AW_DEV int xt_site_kind(const View<T> &V, int i) { return (V.rinfo[i] & RI_GLY) ? K_CA : K_CB; }

// compute_membrane_potential (fb.cpp:6107-6235), the cylindrical pore of
// memb_pore_type 0 (the other form is commented out in the reference)
template <class T>
AW_DEV void xt_memb(const View<T> &V, const int *zres, int i, XtAcc &A)
{
  ParamPtr P = V.P;
  int kind = xt_site_kind(V, i);
  const double *x = XA(V, kind, i);
  double dx = x[0] - P->memb_xo[0], dy = x[1] - P->memb_xo[1], dz = x[2] - P->memb_xo[2];
  double kb = P->memb_kbin, a = P->memb_rho0_distor * P->memb_rho0_max, b = P->memb_len / 2;
  double rho = sqrt(dx * dx + dy * dy);
  double rho0 = (P->memb_rho0_max - a) + (a / P->memb_len) * (dz + b);
  double tper = tanh(kb * (dz - b)), tm1 = tanh(kb * (dz + b)), tm2 = tanh(kb * (b - dz));
  double tcyt = tanh(kb * (-b - dz)), tpor = tanh(kb * (rho - rho0));
  double s_per = 0.5 * (1 + tper), s_mem = 0.5 * (tm1 + tm2), s_cyt = 0.5 * (1 + tcyt);
  double s_por = 0.5 * (1 - tpor), s_nopor = 1 - s_por;
  int z = zres[i] - 1;
  const auto &g = P->g_memb[z];
  // classes 1 and 3 share the signs of fb.cpp:6163-6172, class 2 reverses them
  double sg = z == 1 ? -1.0 : 1.0;
  double Vm = sg * (-g[0] * s_per - g[1] * s_cyt + g[2] * s_mem * s_nopor - g[3] * s_mem * s_por);
  double ek = P->epsilon * P->k_memb;
  A.e[XT_MEMB] += ek * Vm;
  double dz_per = 0.5 * kb * (1 - tper * tper);
  double dz_mem = -0.5 * kb * tm1 * tm1 + 0.5 * kb * tm2 * tm2;
  double dz_cyt = -0.5 * kb * (1 - tcyt * tcyt);
  double sech2 = 1 - tpor * tpor;
  double dz_por = 0.5 * kb * sech2 * (a / P->memb_len);
  double dx_por = ((-0.5 * kb * dx) / rho) * sech2, dy_por = ((-0.5 * kb * dy) / rho) * sech2;
  double dz_por_mem = s_mem * dz_por + dz_mem * s_por;
  double dz_nopor_mem = -s_mem * dz_por + dz_mem * s_nopor;
  double dvx = sg * (-g[2] * s_mem * dx_por - g[3] * s_mem * dx_por);
  double dvy = sg * (-g[2] * s_mem * dy_por - g[3] * s_mem * dy_por);
  double dvz = sg * (-g[0] * dz_per - g[1] * dz_cyt + g[2] * dz_nopor_mem - g[3] * dz_por_mem);
  double *p = FA(V, kind, i);
  aadd(p + 0, -ek * dvx);
  aadd(p + 1, -ek * dvy);
  aadd(p + 2, -ek * dvz);
}

// compute_vector_fragment_memory_potential (fb.cpp:3994-4077), one
// (memory, i, j) item: the angle between the CA->CB vectors of i and j
// against the memory's (it.a)
template <class T>
AW_DEV void xt_vfm(const View<T> &V, const XtPair &it, XtAcc &A)
{
  ParamPtr P = V.P;
  int i = it.i, j = it.j;
  const double *ai = XA(V, K_CA, i), *bi = XA(V, K_CB, i), *aj = XA(V, K_CA, j), *bj = XA(V, K_CB, j);
  double vi[3] = {bi[0] - ai[0], bi[1] - ai[1], bi[2] - ai[2]}, vj[3] = {bj[0] - aj[0], bj[1] - aj[1], bj[2] - aj[2]};
  double vmsqi = vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2];
  double vmsqj = vj[0] * vj[0] + vj[1] * vj[1] + vj[2] * vj[2];
  double vmi = sqrt(vmsqi), vmj = sqrt(vmsqj), vp = vi[0] * vj[0] + vi[1] * vj[1] + vi[2] * vj[2];
  double vpn = vp / (vmi * vmj);
  double dg = acos(vpn) - it.a;
  double Vv = -P->epsilon * P->k_vfm * exp(-dg * dg / (2 * P->vfm_sigma_sq));
  A.e[XT_VFM] += Vv;
  double force = -Vv * dg / (P->vfm_sigma_sq * vmi * vmj * sqrt(1 - vpn * vpn));
  double fi[3], fj[3];
  for (int k = 0; k < 3; ++k) {
    fi[k] = force * (vj[k] - vi[k] * vp / vmsqi);
    fj[k] = force * (vi[k] - vj[k] * vp / vmsqj);
  }
  xt_fadd(V, K_CA, i, -1.0, fi); xt_fadd(V, K_CB, i, 1.0, fi);
  xt_fadd(V, K_CA, j, -1.0, fj); xt_fadd(V, K_CB, j, 1.0, fj);
}

// compute_contact_restraints_potential (fb.cpp:8137-8177): a pair of
// density sites (the host keeps pairs of other chains or |i - j| > 1,
// fb.cpp:8059) within the global cutoff, V = -w exp(-dr^2 / (2 sigma^2))
// for |dr| < 4.29 sigma
template <class T>
AW_DEV void xt_cr(const View<T> &V, const XtPair &it, double cut2, XtAcc &A)
{
  ParamPtr P = V.P;
  int ki = xt_site_kind(V, it.i), kj = xt_site_kind(V, it.j);
  double d[3];
  double rsq = xt_site_sq(XA(V, ki, it.i), XA(V, kj, it.j), d);
  if (!(rsq < cut2)) return;
  double r = sqrt(rsq), dr = r - it.b;
  if (!(fabs(dr) < P->cr_dr_cutoff)) return;
  double Vc = -it.a * exp(-0.5 * dr * dr * P->cr_sigma_sq_inv);
  A.e[XT_CR] += Vc;
  double force = Vc * dr * P->cr_sigma_sq_inv / r;
  xt_fadd(V, ki, it.i, force, d);
  xt_fadd(V, kj, it.j, -force, d);
}

// compute_DebyeHuckel_Interaction (fb.cpp:6307-6363) for a charged pair
// (the host keeps |i - j| > 1 and >= the minimum separation; it.a = k q_i
// q_j with k by the charges' signs) whose CAs are within the list
template <class T>
AW_DEV void xt_dh(const View<T> &V, const XtPair &it, double lr2, XtAcc &A)
{
  ParamPtr P = V.P;
  double d[3];
  if (!(xt_site_sq(XA(V, K_CA, it.i), XA(V, K_CA, it.j), d) < lr2)) return;
  int ki = xt_site_kind(V, it.i), kj = xt_site_kind(V, it.j);
  double r = sqrt(xt_site_sq(XA(V, ki, it.i), XA(V, kj, it.j), d));
  double E = P->epsilon * (it.a / r) * exp(-P->dh_kscreen * r / P->dh_length);
  A.e[XT_DH] += E;
  double ft = (E / r) * (1.0 / r + P->dh_kscreen / P->dh_length);
  xt_fadd(V, ki, it.i, ft, d);
  xt_fadd(V, kj, it.j, -ft, d);
}

// compute_solvent_barrier (fb.cpp:6237-6305) for one residue pair i < j of
// the pair sweep: another chain or j - i >= ssb_ij_sep, CAs within the list
template <class T>
AW_DEV void xt_ssb(const View<T> &V, int i, int j, double lr2, XtAcc &A)
{
  ParamPtr P = V.P;
  int ri = V.rinfo[i], rj = V.rinfo[j];
  if (rchain(ri) == rchain(rj) && j - i < P->ssb_ij_sep) return;
  double d[3];
  if (!(xt_site_sq(XA(V, K_CA, i), XA(V, K_CA, j), d) < lr2)) return;
  int ki = (ri & RI_GLY) ? K_CA : K_CB, kj = (rj & RI_GLY) ? K_CA : K_CB;
  double r = sqrt(xt_site_sq(XA(V, ki, i), XA(V, kj, j), d));
  double rmin1 = P->ssb_rmin1, rmax1 = P->ssb_rmax1, rmin2 = P->ssb_rmin2, rmax2 = P->ssb_rmax2;
  if (P->ssb_rad_cor) {
    double sh = P->ssb_rshift[rtype(ri)] + P->ssb_rshift[rtype(rj)];
    rmin1 += sh; rmax1 += sh; rmin2 += sh; rmax2 += sh;
  }
  double kp = P->ssb_kappa;
  if (r > rmax1 + 10 / kp && r > rmax2 + 10 / kp) return;
  double tmin1 = tanh(kp * (r - rmin1)), tmax1 = tanh(kp * (rmax1 - r));
  double tmin2 = tanh(kp * (r - rmin2)), tmax2 = tanh(kp * (rmax2 - r));
  double th1 = 0.5 * (tmin1 + tmax1), th2 = 0.5 * (tmin2 + tmax2), e = P->epsilon;
  A.e[XT_SSB] += e * P->k_ssb1 * th1 + e * P->k_ssb2 * th2;
  double f = -e * P->k_ssb1 * kp * th1 * (tmax1 - tmin1) / r - e * P->k_ssb2 * kp * th2 * (tmax2 - tmin2) / r;
  xt_fadd(V, ki, i, f, d);
  xt_fadd(V, kj, j, -f, d);
}

// compute_amh_go_model (fb.cpp:3854-3992) for one atom a = 2 i + k (CA,
// or CB of a non-glycine): E_a = sum over a's items (native contacts the
// host kept: rnative < rc, not censored, |i - j| >= minSep or another
// chain) inside the list and (r - rn)^2 < 27.6 sigma^2 of gamma exp(-(r -
// rn)^2 / (2 sigma^2)); the items' forces times -k p E_a^(p-1) / (2 a_norm);
// energy -k E_a^p / (2 a_norm)
template <class T>
AW_DEV void xt_go(const View<T> &V, const int *start, const XtGo *items, double norm, double lr2, int a, XtAcc &A)
{
  ParamPtr P = V.P;
  int i = a >> 1, ka = a & 1;
  const double *xa = XA(V, ka, i);
  int k0 = start[a], k1 = start[a + 1];
  double Ea = 0.0;
  for (int k = k0; k < k1; ++k) {
    const XtGo &it = items[k];
    double d[3];
    double rsq = xt_site_sq(xa, XA(V, it.b & 1, it.b >> 1), d);
    if (!(rsq < lr2)) continue;
    double dr = sqrt(rsq) - it.rn;
    if (!(dr * dr < 27.6 * it.s2)) continue;
    Ea += it.gamma * exp(-dr * dr / (2.0 * it.s2));
  }
  double factor = -0.5 * P->k_amh_go * P->amh_go_p * pow(Ea, P->amh_go_p - 1) / norm;
  double fs[3] = {0.0, 0.0, 0.0};
  for (int k = k0; k < k1; ++k) {
    const XtGo &it = items[k];
    double d[3];
    double rsq = xt_site_sq(xa, XA(V, it.b & 1, it.b >> 1), d);
    if (!(rsq < lr2)) continue;
    double r = sqrt(rsq), dr = r - it.rn;
    if (!(dr * dr < 27.6 * it.s2)) continue;
    double force = factor * it.gamma * exp(-dr * dr / (2.0 * it.s2)) * dr / (it.s2 * r);
    fs[0] += force * d[0]; fs[1] += force * d[1]; fs[2] += force * d[2];
    xt_fadd(V, it.b & 1, it.b >> 1, -force, d);
  }
  double *p = FA(V, ka, i);
  aadd(p + 0, fs[0]);
  aadd(p + 1, fs[1]);
  aadd(p + 2, fs[2]);
  A.e[XT_GO] += -0.5 * P->k_amh_go * pow(Ea, P->amh_go_p) / norm;
}

// every coeff-toggled term of one replica by the lanes t0, t0 + ts, ...
// (after the list build: the barrier sweeps the Verlet list when valid);
// energies into ered[N_EACC + k] when reported (ef)
template <class T>
AW_NOINLINE void xt_pass(const View<T> &V, const ReplicaDesc &R, int t0, int ts, bool ef)
{
  ParamPtr P = V.P;
  if (!P->xt_on) return;
  XtAcc A;
  for (int k = 0; k < N_XT; ++k) A.e[k] = 0.0;
  int n = V.n;
  const double lr2 = R.list_r2;
  if (P->memb_flag && ON(P, AW_ET_MEMB) && R.zres)
    for (int i = t0; i < n; i += ts) xt_memb(V, R.zres, i, A);
  if (P->vfm_flag && ON(P, AW_ET_VFRAGMEM)) {
    const XtPair *it = (const XtPair *)R.vfm_item;
    for (int k = t0; k < R.n_vfm; k += ts) xt_vfm(V, it[k], A);
  }
  if (P->cr_flag && ON(P, AW_ET_CONT_REST)) {
    const XtPair *it = (const XtPair *)R.cr_item;
    const double c2 = R.cr_cut2;
    for (int k = t0; k < R.n_cr; k += ts) xt_cr(V, it[k], c2, A);
  }
  if (P->dh_flag && ON(P, AW_ET_DH)) {
    const XtPair *it = (const XtPair *)R.dh_item;
    for (int k = t0; k < R.n_dh; k += ts) xt_dh(V, it[k], lr2, A);
  }
  if (P->amhgo_flag && ON(P, AW_ET_AMHGO) && R.go_start)
    for (int a = t0; a < 2 * n; a += ts)
      if (!((a & 1) && (V.rinfo[a >> 1] & RI_GLY))) xt_go(V, R.go_start, (const XtGo *)R.go_item, R.go_norm, lr2, a, A);
  if (P->ssb_flag && ON(P, AW_ET_SSB)) for_pairs(V, t0, ts, [&](int i, int j) { xt_ssb(V, i, j, lr2, A); });
  if (!ef) return;
#ifdef AW_HOST_EMU
  for (int k = 0; k < N_XT; ++k) V.ered[N_EACC + k] += A.e[k];
#else
  int lane = threadIdx.x & 63;
  for (int k = 0; k < N_XT; ++k) {
    double v = wave_sum_d(A.e[k]);
    if (lane == 0 && v != 0.0) atomicAdd(&V.ered[N_EACC + k], v);
  }
#endif
}
