// ref_shim.cpp -- C wrapper over the reference's standalone sources, used to
// pin the oracle (test infrastructure only; built into oracle/_ref/).
//   src/fragment_memory.cpp : Fragment_Memory (.gro reader, Rf), Gamma_Array
//   src/smart_matrix_lib.h  : cWell::compute_theta_pair, cP_AP::compute, cR
// The templates in smart_matrix_lib.h read geometry through their second
// template argument; ShimHost below supplies exactly the members they use.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include "fragment_memory.h"
#include "smart_matrix_lib.h"

struct ShimHost {
  enum ResInfo { NONE = 0, LOCAL, GHOST, OFF };
  double *xca[2], *xcb[2], *xo[2], *xn[2], *xh[2];
  double buf[10][3];
  double P_AP_cut, P_AP_pref;
  int nn, res_no[2], res_info[2], chain_no[2];
  char se[3];
  ShimHost() {
    for (int k = 0; k < 2; ++k) {
      xca[k] = buf[k]; xcb[k] = buf[2 + k]; xo[k] = buf[4 + k]; xn[k] = buf[6 + k]; xh[k] = buf[8 + k];
      res_no[k] = k + 1; res_info[k] = LOCAL; chain_no[k] = 1;
    }
    nn = 2; se[0] = se[1] = 'A'; se[2] = 0;
    P_AP_cut = 0; P_AP_pref = 1;
  }
};

extern "C" {

void *ref_gamma_create(const char *path) { return new Gamma_Array(const_cast<char *>(path)); }
int ref_gamma_error(void *h) { return static_cast<Gamma_Array *>(h)->error; }
double ref_gamma_get2(void *h, int it, int jt, int i, int j) { return static_cast<Gamma_Array *>(h)->getGamma(it, jt, i, j); }
double ref_gamma_get4(void *h, int it, int jt, int ift, int jft, int i, int j) { return static_cast<Gamma_Array *>(h)->getGamma(it, jt, ift, jft, i, j); }
int ref_gamma_minsep(void *h) { return static_cast<Gamma_Array *>(h)->minSep(); }
int ref_gamma_maxsep(void *h) { return static_cast<Gamma_Array *>(h)->maxSep(); }
int ref_gamma_four(void *h) { return static_cast<Gamma_Array *>(h)->fourResTypes() ? 1 : 0; }
void ref_gamma_destroy(void *h) { delete static_cast<Gamma_Array *>(h); }

void *ref_mem_create(int pos, int fpos, int len, double w, const char *path) {
  return new Fragment_Memory(pos, fpos, len, w, const_cast<char *>(path));
}
int ref_mem_error(void *h) { return static_cast<Fragment_Memory *>(h)->error; }
double ref_mem_rf(void *h, int ires, int iatom, int jres, int jatom) {
  return static_cast<Fragment_Memory *>(h)->Rf(ires, iatom, jres, jatom);
}
char ref_mem_se(void *h, int resno) { return static_cast<Fragment_Memory *>(h)->getSe(resno); }
void ref_mem_destroy(void *h) { delete static_cast<Fragment_Memory *>(h); }

void ref_theta_pair(double kappa, double rmin, double rmax, double r, double *theta, double *prd, double *rmin_th, double *rmax_th) {
  int flag[1] = {1};
  double rmn[1] = {rmin}, rmx[1] = {rmax};
  WPV par(kappa, 1.0, 0.0, 1, flag, rmn, rmx);
  ShimHost host;
  int step = 1;
  cWell<double, ShimHost> well(2, 2, 1, par, &step, &host);
  *theta = well.theta_pair(0, 1, 0, r);
  *prd = well.prd_theta_pair(0, 1, 0, r);
  *rmin_th = well.rmin_theta[0];
  *rmax_th = well.rmax_theta[0];
}

void ref_pap_nu(double cut, double pref, const double *xi, const double *xj, double *nu, double *prd) {
  ShimHost host;
  host.P_AP_cut = cut; host.P_AP_pref = pref;
  for (int k = 0; k < 3; ++k) { host.xca[0][k] = xi[k]; host.xca[1][k] = xj[k]; }
  int step = 1;
  cP_AP<double, ShimHost> p(2, 2, &step, &host);
  *nu = p.nu(0, 1);
  *prd = p.prd_nu(0, 1);
}

void ref_rno_rho(const double *o, const double *n, const double *h, double *rno, double *rho) {
  ShimHost host;
  for (int k = 0; k < 3; ++k) { host.xo[0][k] = o[k]; host.xn[1][k] = n[k]; host.xh[1][k] = h[k]; }
  int step = 1;
  cR<double, ShimHost> R(2, 2, &step, &host);
  *rno = R.rNO(0, 1);
  *rho = R.rHO(0, 1);
}

}  // extern "C"
