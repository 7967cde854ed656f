// awsem_parse.cpp -- readers for the reference's input formats.
//
// Semantics follow the FixBackbone constructor (src/fix_backbone.cpp):
// the coeff file is a stream of whitespace tokens where a token equal to a
// section tag ("[Water]") starts that section and anything else is skipped
// (a "-" suffix on a tag therefore disables a section, fb.cpp:165-497);
// numeric fields use istream extraction.  Side files are read from the
// working directory by their fixed names (fb.cpp:555-693); fragment
// memories use the .mem / .gro formats (fb.cpp:1966-2042,
// src/fragment_memory.cpp:34-176) and the gamma-class file format
// (src/fragment_memory.cpp:275-378).
#include "awsem_host.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>
#include <map>
#include <algorithm>

static const int kSeMap[26] = {0, 0, 4, 3, 6, 13, 7, 8, 9, 0, 11, 10, 12, 2, 0, 14, 5, 1, 15, 16, 0, 19, 17, 0, 18, 0};
static const int kFour[20] = {1, 3, 2, 2, 4, 2, 2, 1, 3, 4, 4, 3, 4, 4, 1, 1, 1, 4, 4, 4};

int aw_restype(char c) { return (c >= 'A' && c <= 'Z') ? kSeMap[c - 'A'] : 0; }

static std::string join(const std::string &dir, const std::string &name)
{
  if (name.empty() || name[0] == '/' || dir.empty()) return name;
  return dir + "/" + name;
}

static const char *kUnsupported[] = {
    "[Amylometer]", "[Selection_Temperature]",
    "[Monte_Carlo_Seq_Opt]", "[Optimization]", "[Burial_Optimization]",
    "[DebyeHuckel_Optimization]", nullptr};

AwStatus aw_parse_coeff(const std::string &path, FFParams &P, std::string &mem_file, std::string &gamma_file,
                        std::string &cr_file, std::string &ff_file)
{
  std::ifstream in(path);
  if (!in) return AwStatus::err(-2, "Coefficient file was not found!");
  std::memset(&P, 0, sizeof(P));
  P.epsilon = 1.0;
  P.an = 0.4831806; P.bn = 0.7032820; P.cn = -0.1864262;   // fb.cpp:141-143
  P.ap = 0.4436538; P.bp = 0.2352006; P.cp = 0.3211455;
  P.ah = 0.8409657; P.bh = 0.8929599; P.ch = -0.7338894;
  P.fm_sigma_exp = 0.15;
  P.term_mask = 0xffffffffu;
  double sigma[12] = {0}, sigma_HO = 0, sigma_NO = 0, hsig_HO = 0, hsig_NO = 0;
  std::string sec;
  while (in >> sec) {
    if (sec == "[ABC]") {
      in >> P.an >> P.bn >> P.cn >> P.ap >> P.bp >> P.cp >> P.ah >> P.bh >> P.ch;
    } else if (sec == "[Chain]") {
      P.chain_flag = 1;
      in >> P.k_chain[0] >> P.k_chain[1] >> P.k_chain[2] >> P.r_ncb0 >> P.r_cpcb0 >> P.r_ncp0;
    } else if (sec == "[Shake]") {
      P.shake_flag = 1;
      in >> P.k_shake >> P.r_sh1 >> P.r_sh2 >> P.r_sh3;
    } else if (sec == "[Chi]") {
      P.chi_flag = 1;
      in >> P.k_chi >> P.chi0;
    } else if (sec == "[Excluded]" || sec == "[Excluded_R6]") {
      double d;
      for (int k = 0; k < 4; ++k) in >> d;  // only used by the DEBUGFORCES build
    } else if (sec == "[Excluded_P]") {
      int p; double d;
      in >> p;
      for (int k = 0; k < 4; ++k) in >> d;
    } else if (sec == "[Rama]") {
      P.rama_flag = 1;
      in >> P.k_rama >> P.n_rama_par;
      if (P.n_rama_par < 0 || P.n_rama_par > 6) return AwStatus::err(-3, "Rama: too many wells");
      for (int j = 0; j < P.n_rama_par; ++j) {
        in >> P.w[j] >> sigma[j] >> P.phiw[j] >> P.phi0[j] >> P.psiw[j] >> P.psi0[j];
        P.phiw[j] *= sigma[j];
        P.psiw[j] *= sigma[j];
      }
    } else if (sec == "[Rama_P]") {
      P.rama_p_flag = 1;
      in >> P.n_rama_p_par;
      if (P.n_rama_p_par < 0 || P.n_rama_p_par > 6) return AwStatus::err(-3, "Rama_P: too many wells");
      for (int j = 6; j < 6 + P.n_rama_p_par; ++j) {
        in >> P.w[j] >> sigma[j] >> P.phiw[j] >> P.phi0[j] >> P.psiw[j] >> P.psi0[j];
        P.phiw[j] *= sigma[j];
        P.psiw[j] *= sigma[j];
      }
    } else if (sec == "[SSWeight]") {
      P.ssweight_flag = 1;
      for (int j = 0; j < 12; ++j) { bool b = false; in >> b; P.ssweight[j] = b ? 1 : 0; }
    } else if (sec == "[Dssp_Hdrgn]") {
      P.dssp_flag = 1;
      in >> P.k_dssp >> P.hbscl[0][0] >> P.hbscl[0][1];
      for (int j = 0; j < 7; ++j) in >> P.hbscl[1][j];
      for (int j = 0; j < 9; ++j) in >> P.hbscl[2][j];
      for (int j = 0; j < 9; ++j) in >> P.hbscl[3][j];
      in >> sigma_HO >> sigma_NO >> P.HO_zero >> P.NO_zero >> P.dssp_cut >> P.pref[0] >> P.pref[1] >> P.d_nu0;
      P.dssp_cut_sq = P.dssp_cut * P.dssp_cut;
      P.nu_cut1_sq = std::pow(P.d_nu0 + std::atanh(2.0 * 1e-4 - 1.0) / P.pref[0], 2);
      P.nu_cut2_sq = std::pow(P.d_nu0 + std::atanh(2.0 * 1e-4 - 1.0) / P.pref[1], 2);
      P.sigma_HO_sqinv = 1.0 / (sigma_HO * sigma_HO);
      P.sigma_NO_sqinv = 1.0 / (sigma_NO * sigma_NO);
    } else if (sec == "[P_AP]") {
      P.pap_flag = 1;
      in >> P.k_global_pap >> P.k_betapred_pap >> P.k_pap[0] >> P.k_pap[1] >> P.k_pap[2] >> P.pap_cut >> P.pap_pref;
      in >> P.i_med_min >> P.i_med_max >> P.i_diff_pap;
      P.pap_cutoff_sq = std::pow(P.pap_cut + std::atanh(1.0 - 2.0 * 1e-12) / P.pap_pref, 2);
    } else if (sec == "[Water]") {
      P.water_flag = 1;
      in >> P.k_water >> P.water_kappa >> P.water_kappa_sigma >> P.treshold >> P.contact_cutoff >> P.n_wells;
      if (P.n_wells < 0 || P.n_wells > 5) return AwStatus::err(-3, "Water: too many wells");
      for (int j = 0; j < P.n_wells; ++j) in >> P.well_r_min[j] >> P.well_r_max[j] >> P.well_flag[j];
    } else if (sec == "[Burial]") {
      P.burial_flag = 1;
      in >> P.k_burial >> P.burial_kappa;
      for (int j = 0; j < 3; ++j) in >> P.burial_ro_min[j] >> P.burial_ro_max[j];
    } else if (sec == "[Helix]") {
      P.helix_flag = 1;
      in >> P.k_helix >> P.helix_gamma_p >> P.helix_gamma_w >> P.helix_kappa >> P.helix_kappa_sigma;
      in >> P.helix_treshold >> P.helix_i_diff >> P.helix_cutoff;
      P.helix_cutoff_sq = P.helix_cutoff * P.helix_cutoff;
      in >> P.helix_r_min >> P.helix_r_max;
      for (int j = 0; j < 20; ++j) in >> P.h4prob[j];
      in >> P.pro_accepter_flag >> P.h4prob_pro_accepter;
      in >> hsig_HO >> hsig_NO >> P.helix_HO_zero >> P.helix_NO_zero;
      P.helix_sigma_HO_sqinv = 1.0 / (hsig_HO * hsig_HO);
      P.helix_sigma_NO_sqinv = 1.0 / (hsig_NO * hsig_NO);
    } else if (sec == "[Fragment_Memory]") {
      P.fm_flag = 1;
      in >> P.k_frag_mem >> mem_file >> gamma_file;
    } else if (sec == "[Fragment_Memory_Table]") {
      // fb.cpp:310-320
      P.fm_tb_flag = 1;
      in >> P.k_frag_mem >> mem_file >> gamma_file >> P.tb_rmin >> P.tb_rmax >> P.tb_dr;
      P.tb_size = (int)((P.tb_rmax - P.tb_rmin) / P.tb_dr) + 2;
      in >> P.tb_well_width >> P.fm_use_table >> P.fm_sigma_exp;
      if (!in.fail() && (!(P.tb_dr > 0.0) || P.tb_size < 2))
        return AwStatus::err(-3, "Fragment_Memory_Table: empty distance grid");
    } else if (sec == "[Epsilon]") {
      in >> P.epsilon;
    } else if (sec == "[AMH-Go]") {  // fb.cpp:297-303
      P.amhgo_flag = 1;
      in >> P.k_amh_go >> P.amh_go_p >> P.amh_go_rc >> P.amh_go_censor;
    } else if (sec == "[Contact_Restraints]") {  // fb.cpp:321-327
      P.cr_flag = 1;
      in >> P.k_cont_rest >> P.cr_sigma >> cr_file;
      P.cr_sigma_sq_inv = std::pow(P.cr_sigma, -2);
    } else if (sec == "[Vector_Fragment_Memory]") {  // fb.cpp:328-333
      P.vfm_flag = 1;
      in >> P.k_vfm >> P.vfm_sigma;
      P.vfm_sigma_sq = P.vfm_sigma * P.vfm_sigma;
    } else if (sec == "[Solvent_Barrier]") {  // fb.cpp:334-345
      P.ssb_flag = 1;
      in >> P.k_ssb1 >> P.ssb_rmin1 >> P.ssb_rmax1 >> P.k_ssb2 >> P.ssb_rmin2 >> P.ssb_rmax2 >> P.ssb_kappa;
      in >> P.ssb_ij_sep >> P.ssb_rad_cor;
      for (int j = 0; j < 20; ++j) in >> P.ssb_rshift[j];
    } else if (sec == "[Membrane]") {  // fb.cpp:346-358
      P.memb_flag = 1;
      in >> P.k_memb >> P.memb_kbin >> P.memb_xo[0] >> P.memb_xo[1] >> P.memb_xo[2] >> P.memb_pore_type;
      in >> P.memb_len >> P.memb_rho0_max >> P.memb_rho0_distor;
      for (int a = 0; a < 3; ++a)
        for (int j = 0; j < 4; ++j) in >> P.g_memb[a][j];
    } else if (sec == "[Phosphorylation]") {  // fb.cpp:412-421
      if (!P.water_flag) return AwStatus::err(-3, "Cannot run phosphorylation without water potential");
      P.phosph_flag = 1;
      in >> P.k_hypercharge >> P.n_phosph;
      if (P.n_phosph > 20) return AwStatus::err(-3, "Number of phosphorylated residues may not exceed 20");
      for (int j = 0; j < P.n_phosph; ++j) in >> P.phosph_res[j];
    } else if (sec == "[Tertiary_Frustratometer]") {  // fb.cpp:385-397
      std::string mode;
      P.tf_flag = 1;
      in >> P.tf_cutoff >> P.tf_ndecoys >> P.tf_freq >> mode;
      if (mode == "configurational") P.tf_mode = 0;
      else if (mode == "mutational") P.tf_mode = 1;
      else if (mode == "singleresidue") P.tf_mode = 2;
      else if (!in.fail())
        return AwStatus::err(-3, "Only \"configurational\", \"mutational\", \"singleresidue\" are acceptable modes "
                                 "for the Tertiary_Frustratometer.");
    } else if (sec == "[Fragment_Frustratometer]") {  // fb.cpp:359-384
      if (!P.fm_flag && !P.fm_tb_flag)
        return AwStatus::err(-3, "Cannot run Fragment_Frustratometer without Fragment_Memory or Fragment_Memory_Table.");
      std::string mode;
      P.ff_flag = 1;
      P.ff_ww = 1.0;
      in >> mode;
      if (mode == "shuffle") {
        P.ff_mode = 0;
        in >> ff_file >> P.ff_ncalc >> P.ff_freq;
      } else if (mode == "read") {
        P.ff_mode = 1;
        in >> ff_file >> P.ff_freq >> P.ff_ww >> P.ff_seqsep >> P.ff_seqsep_gamma >> P.ff_norm;
      } else {
        return AwStatus::err(-3, "Only \"shuffle\" and \"read\" are acceptable modes for the Fragment_Frustratometer.");
      }
    } else if (sec == "[Nmer_Frustratometer]") {  // fb.cpp:398-412
      std::string mode;
      P.nf_flag = 1;
      in >> P.nf_size >> P.nf_cutoff >> P.nf_contacts >> P.nf_ndecoys >> P.nf_freq;
      in >> P.nf_min_thr >> P.nf_high_thr >> P.nf_neutral;
      in >> P.nf_trap >> P.nf_draw_trap >> P.nf_trap_sigma >> P.nf_ss_frac >> mode;
      if (mode == "pairwise") P.nf_mode = 0;
      else if (mode == "singlenmer") P.nf_mode = 1;
      else if (!in.fail())
        return AwStatus::err(-3, "Only \"pairwise\", \"singlenmer\" are acceptable modes for the Nmer_Frustratometer.");
    } else if (sec == "[DebyeHuckel]") {  // fb.cpp:466-478
      P.dh_flag = 1;
      in >> P.dh_k[0] >> P.dh_k[1] >> P.dh_k[2] >> P.dh_kscreen >> P.dh_length >> P.dh_min_sep;
    } else if (sec == "[Shuffler]" || sec == "[Mutate_Sequence]") {
      int fl = 0;
      std::string name;
      in >> fl >> name;
      if (fl == 1) return AwStatus::err(-4, "Section " + sec + " is enabled but is not part of this build's force path");
    } else {
      for (int k = 0; kUnsupported[k]; ++k)
        if (sec == kUnsupported[k])
          return AwStatus::err(-4, "Section " + sec + " is enabled but is not part of this build's force path");
    }
    if (in.fail()) return AwStatus::err(-3, "Coefficient file format error in section " + sec);
  }
  // fb.cpp:502-516: scale strengths by epsilon
  double e = P.epsilon;
  for (int j = 0; j < 3; ++j) P.k_chain[j] *= e;
  P.k_chi *= e;
  P.k_rama *= e;
  P.k_water *= e;
  P.k_burial *= e;
  P.k_helix *= e;
  P.k_dssp *= e;
  P.k_global_pap *= e;
  P.k_amh_go *= e;
  P.k_cont_rest *= e;
  for (int j = 0; j < P.n_rama_par; ++j) P.w[j] *= P.k_rama;
  for (int j = 6; j < 6 + P.n_rama_p_par; ++j) P.w[j] *= P.k_rama;
  for (int j = 0; j < 12; ++j) {
    P.cphi0[j] = std::cos(P.phi0[j]); P.sphi0[j] = std::sin(P.phi0[j]);
    P.cpsi0[j] = std::cos(P.psi0[j]); P.spsi0[j] = std::sin(P.psi0[j]);
  }
  // cWell ranges (sml.h:363-368), cP_AP ranges (sml.h:111-115)
  for (int k = 0; k < P.n_wells; ++k) {
    P.w_rmin_th[k] = P.well_r_min[k] - 8.0 * 2.302585 / P.water_kappa;
    P.w_rmax_th[k] = P.well_r_max[k] + 8.0 * 2.302585 / P.water_kappa;
    P.w_rmin_th_sq[k] = P.w_rmin_th[k] * P.w_rmin_th[k];
    P.w_rmax_th_sq[k] = P.w_rmax_th[k] * P.w_rmax_th[k];
  }
  if (P.helix_flag) {
    P.h_rmin_th = P.helix_r_min - 8.0 * 2.302585 / P.helix_kappa;
    P.h_rmax_th = P.helix_r_max + 8.0 * 2.302585 / P.helix_kappa;
    P.h_rmin_th_sq = P.h_rmin_th * P.h_rmin_th;
    P.h_rmax_th_sq = P.h_rmax_th * P.h_rmax_th;
  }
  if (P.pap_flag) {
    double dd = (double)((int)(8.0 * 2.302585 / P.pap_pref) + 1);
    P.pap_drmax_sq = (P.pap_cut + dd) * (P.pap_cut + dd);
    P.pap_drmin_sq = (P.pap_cut - dd) * (P.pap_cut - dd);
  }
  // pair_list_cutoff, fb.cpp:1048-1104
  double plc = 8.0;
  if (P.dssp_flag) plc = std::fmax(plc, P.dssp_cut);
  if (P.pap_flag) plc = std::fmax(plc, std::sqrt(P.pap_cutoff_sq));
  if (P.water_flag) for (int i = 0; i < P.n_wells; ++i) plc = std::fmax(plc, P.w_rmax_th[i]);
  if (P.helix_flag) {
    plc = std::fmax(plc, P.h_rmax_th);
    double r0[3] = {3.77, 2.5, 2.41};
    auto three = [&](double a, double b, double c) {
      double r[3];
      r[0] = std::sqrt(std::pow(a * r0[0], 2) + std::pow(c * r0[1], 2) + a * c * (r0[0] * r0[0] + r0[1] * r0[1] - r0[2] * r0[2]));
      r[1] = std::sqrt(std::pow(a * r0[2], 2) + std::pow(b * r0[1], 2) + a * b * (r0[2] * r0[2] + r0[1] * r0[1] - r0[0] * r0[0]));
      r[2] = std::sqrt(std::pow(b * r0[0], 2) + std::pow(c * r0[2], 2) + b * c * (r0[0] * r0[0] + r0[2] * r0[2] - r0[1] * r0[1]));
      return std::fmax(r[0], std::fmax(r[1], r[2]));
    };
    double r1 = 1.1 * three(P.an, P.bn, P.cn) + P.helix_NO_zero + hsig_NO * std::sqrt(6.0 * std::log(10.0));
    double r2 = 1.1 * three(P.ah, P.bh, P.ch) + P.helix_HO_zero + hsig_HO * std::sqrt(6.0 * std::log(10.0));
    plc = std::fmax(plc, std::fmax(r1, r2));
    plc = std::fmax(plc, P.helix_cutoff);
  }
  // (the contact restraints' share follows their file, aw_read_extras)
  if (P.dh_flag) plc = std::fmax(plc, 2.0 * P.dh_length / P.dh_kscreen);
  if (P.ssb_flag) plc = std::fmax(plc, aw_ssb_cut(P));
  P.pair_list_cutoff = plc;
  P.xt_on = (P.amhgo_flag || P.cr_flag || P.vfm_flag || P.ssb_flag || P.memb_flag || P.dh_flag) ? 1 : 0;
  aw_default_pair_styles(P);
  P.nl_skin = 2.0;
  return AwStatus();
}

// the barrier's reach (fb.cpp:1060-1066): twice the largest radius shift,
// plus the larger outer radius and 10 / kappa
double aw_ssb_cut(const FFParams &P)
{
  double cut = 0.0;
  if (P.ssb_rad_cor)
    for (int i = 0; i < 20; ++i) cut = std::fmax(cut, 2.0 * P.ssb_rshift[i]);
  return cut + std::fmax(P.ssb_rmax1, P.ssb_rmax2) + 10.0 / P.ssb_kappa;
}

void aw_default_pair_styles(FFParams &P)
{
  // excluded volume: pair_style vexcluded 2 3.5 3.5; pair_coeff * * 0.0;
  // 1 1 / 1 4 / 4 4: 20.0 3.5 4.5; 3 3: 20.0 3.5 3.5
  // (tools/create_project_tools/CoordinatesToWorkLammpsDataFile.py:314-322)
  P.ev_on = 1;
  P.ev_p = 2;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) { P.ev_lambda[a][b] = 0.0; P.ev_cut_short[a][b] = 3.5; P.ev_cut_long[a][b] = 3.5; }
  const int tp[4][2] = {{1, 1}, {1, 4}, {4, 4}, {3, 3}};
  const double cl[4] = {4.5, 4.5, 4.5, 3.5};
  for (int k = 0; k < 4; ++k) {
    int a = tp[k][0], b = tp[k][1];
    P.ev_lambda[a][b] = P.ev_lambda[b][a] = 20.0;
    P.ev_cut_long[a][b] = P.ev_cut_long[b][a] = cl[k];
  }
  // bond_style harmonic, K = 200 (CoordinatesToWorkLammpsDataFile.py:289-294)
  P.bond_on = 1;
  P.bond_k = 200.0;
  P.bond_r0[0] = 3.77; P.bond_r0[1] = 2.41; P.bond_r0[2] = 2.50; P.bond_r0[3] = 1.54; P.bond_r0[4] = 1.54;
}

static AwStatus read_numbers(const std::string &path, const char *name, double *dst, int count)
{
  std::ifstream in(path);
  if (!in) return AwStatus::err(-2, std::string("File ") + name + " doesn't exist");
  for (int k = 0; k < count; ++k) {
    if (!(in >> dst[k])) return AwStatus::err(-3, std::string(name) + ": file format error");
  }
  return AwStatus();
}

AwStatus aw_read_param_tables(const std::string &dir, AwSystem &S)
{
  FFParams &P = S.P;
  if (P.dssp_flag) {  // fb.cpp:555-593
    std::vector<double> t(800);
    const char *names[3] = {"anti_HB", "anti_NHB", "para_HB"};
    double(*dst[3])[20][2] = {P.anti_HB, P.anti_NHB, P.para_HB};
    for (int f = 0; f < 3; ++f) {
      AwStatus s = read_numbers(join(dir, names[f]), names[f], t.data(), 800);
      if (!s.ok()) return s;
      for (int i = 0; i < 20; ++i)
        for (int j = 0; j < 20; ++j) { dst[f][i][j][0] = t[i * 20 + j]; dst[f][i][j][1] = t[400 + i * 20 + j]; }
    }
    AwStatus s = read_numbers(join(dir, "para_one"), "para_one", P.para_one, 20);
    if (!s.ok()) return s;
    s = read_numbers(join(dir, "anti_one"), "anti_one", P.anti_one, 20);
    if (!s.ok()) return s;
  }
  double k_water = P.k_water;
  if (P.water_flag) {  // fb.cpp:623-641
    std::ifstream in(join(dir, "gamma.dat"));
    if (!in) return AwStatus::err(-2, "File gamma.dat doesn't exist");
    for (int w = 0; w < P.n_wells; ++w)
      for (int i = 0; i < 20; ++i)
        for (int j = i; j < 20; ++j) {
          double g0, g1;
          if (!(in >> g0 >> g1)) return AwStatus::err(-3, "Water potential gamma.dat file format error");
          P.water_gamma[w][i][j][0] = P.water_gamma[w][j][i][0] = g0 * k_water;
          P.water_gamma[w][i][j][1] = P.water_gamma[w][j][i][1] = g1 * k_water;
        }
  }
  if (P.phosph_flag) {  // fb.cpp:643-669
    for (int w = 0; w < P.n_wells; ++w)
      for (int i = 0; i < 20; ++i)
        for (int j = 0; j < 20; ++j)
          for (int k = 0; k < 2; ++k) P.phosph_gamma[w][i][j][k] = P.water_gamma[w][i][j][k];
    // serine (15) takes glutamate's (6) entries times k, k^2 or 1 by the
    // four-letter class, row by row in this order: an entry written earlier
    // in the loop is read later ((15, 15) from (15, 6), set at i = 6)
    for (int w = 0; w < P.n_wells; ++w)
      for (int i = 0; i < 20; ++i) {
        double f = kFour[i] == 1 ? P.k_hypercharge
                 : (kFour[i] == 2 || kFour[i] == 3) ? std::pow(P.k_hypercharge, 2) : 1.0;
        for (int k = 0; k < 2; ++k)
          P.phosph_gamma[w][i][15][k] = P.phosph_gamma[w][15][i][k] = P.phosph_gamma[w][i][6][k] * f;
      }
  }
  if (P.burial_flag) {  // fb.cpp:684-693
    if (!P.water_flag) return AwStatus::err(-3, "Cannot use burial without water potential");
    double t[60];
    AwStatus s = read_numbers(join(dir, "burial_gamma.dat"), "burial_gamma.dat", t, 60);
    if (!s.ok()) return AwStatus::err(s.code, s.code == -2 ? s.msg : "Burial potential burial_gamma.dat file format error");
    for (int i = 0; i < 20; ++i) for (int k = 0; k < 3; ++k) P.burial_gamma[i][k] = t[i * 3 + k];
  }
  return AwStatus();
}

AwStatus aw_read_sequence(const std::string &path, AwSystem &S)
{
  std::ifstream in(path);
  if (!in) return AwStatus::err(-2, "Sequence file was not found");
  std::string tok;
  S.se.clear(); S.ch_start.clear(); S.ch_len.clear();
  while (in >> tok) {  // fb.cpp:538-553: one whitespace token per chain
    if (tok.empty() || tok[0] == '#') continue;
    for (char c : tok)
      if (c < 'A' || c > 'Z') return AwStatus::err(-3, "Sequence file: unknown residue letter");
    if ((int)S.ch_len.size() >= AW_MAXCH) return AwStatus::err(-3, "Too many chains");
    S.ch_start.push_back((int)S.se.size());
    S.ch_len.push_back((int)tok.size());
    S.se += tok;
  }
  S.nch = (int)S.ch_len.size();
  S.n = (int)S.se.size();
  if (S.n == 0) return AwStatus::err(-3, "Empty sequence");
  return AwStatus();
}

AwStatus aw_read_ssweight(const std::string &dir, AwSystem &S)
{
  S.aps.assign((size_t)S.n * 12, 0.0);
  if (!S.P.ssweight_flag) return AwStatus();
  std::ifstream in(join(dir, "ssweight"));  // fb.cpp:595-604
  if (!in) return AwStatus::err(-2, "File ssweight doesn't exist");
  for (int j = 0; j < S.n; ++j)
    for (int i = 0; i < 12; ++i)
      if (S.P.ssweight[i]) { double v = 0.0; in >> v; S.aps[(size_t)j * 12 + i] = v; }
  return AwStatus();
}

// ---------------------------------------------------------------------------
// Gamma_Array

double AwGamma::get(int it, int jt, int ires, int jres) const
{
  int dij = std::abs(ires - jres);
  if (dij < i_sep[0] || (i_sep[nseq_cl] != -1 && dij > i_sep[nseq_cl])) return 0.0;
  int cl = 1;
  while (cl < nseq_cl && dij >= i_sep[cl]) cl++;
  int r = nres_cl;
  if (r == 1) return gamma[cl - 1];
  int a = it, b = jt;
  if (r == 4) { a = kFour[it] - 1; b = kFour[jt] - 1; }
  return gamma[(size_t)(cl - 1) * r * r + a * r + b];
}

static bool blank(const char *s)
{
  for (; *s; ++s) if (*s != ' ' && *s != '\t' && *s != '\n') return false;
  return true;
}

static int class_indices(const AwGamma &g, const std::string &t, int *a)
{
  if (t == "ALL") { for (int i = 0; i < g.nres_cl; ++i) a[i] = i; return g.nres_cl; }
  if (t.size() == 1) {
    int code = t[0] - 'A';
    if (code < 0 || code >= 26 || (kSeMap[code] == 0 && t[0] != 'A')) return -1;
    if (g.nres_cl < 20) return -2;
    a[0] = kSeMap[code];
    return 1;
  }
  if (t.size() == 3) {
    if (g.nres_cl < 4) return -2;
    int cl = t == "SHL" ? 1 : t == "AHL" ? 2 : t == "BAS" ? 3 : t == "HPB" ? 4 : 0;
    if (!cl) return -1;
    if (g.nres_cl == 4) { a[0] = cl - 1; return 1; }
    int n = 0;
    for (int i = 0; i < 20; ++i) if (kFour[i] == cl) a[n++] = i;
    return n;
  }
  return -1;
}

AwStatus aw_read_gamma_array(const std::string &path, AwGamma &g)
{
  g = AwGamma();
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return AwStatus::err(-2, "Fragment_Memory: Cannot read gamma file");
  char line[100];
  std::vector<std::vector<std::string>> rows;
  int iline = 0;
  bool set = false;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#' || blank(line)) continue;
    if (iline == 0) {
      int ns = 0;
      char *st = std::strtok(line, " \t");
      while (st) {
        ns++;
        if (ns > 10) break;
        if (std::strcmp(st, "inf") == 0) { g.i_sep[ns - 1] = -1; break; }
        g.i_sep[ns - 1] = std::atoi(st);
        st = std::strtok(nullptr, " \t\n");
      }
      g.nseq_cl = ns - 1;
      if (g.nseq_cl <= 0) { std::fclose(f); return AwStatus::err(-3, "Fragment_Memory: Wrong definition of sequance separation classes"); }
    } else {
      std::vector<std::string> tk;
      for (char *st = std::strtok(line, " \t\n"); st; st = std::strtok(nullptr, " \t\n")) tk.push_back(st);
      if (tk.size() != 4 && tk.size() != 6) { std::fclose(f); return AwStatus::err(-3, "Fragment_Memory: Incorrect entery in gamma file"); }
      bool six = tk.size() == 6;
      if (!set) { g.frag_resty = six; set = true; }
      else if (six != g.frag_resty) { std::fclose(f); return AwStatus::err(-3, "Fragment_Memory: Incorrect entery in gamma file"); }
      for (size_t i = 0; i + 2 < tk.size(); ++i) {
        size_t L = tk[i].size();
        if (L != 1 && L != 3) { std::fclose(f); return AwStatus::err(-3, "Fragment_Memory: Incorrect entery in gamma file"); }
        if (tk[i] == "ALL") g.nres_cl = std::max(g.nres_cl, 1);
        else if (L == 3) g.nres_cl = std::max(g.nres_cl, 4);
        else g.nres_cl = std::max(g.nres_cl, 20);
      }
      rows.push_back(tk);
    }
    iline++;
  }
  std::fclose(f);
  if (g.nres_cl == 0) return AwStatus::err(-3, "Fragment_Memory: Incorrect entery in gamma file");
  if (g.frag_resty)
    return AwStatus::err(-4, "Fragment_Memory: gamma files with fragment residue classes are not part of this build's force path");
  int r = g.nres_cl;
  g.gamma.assign((size_t)g.nseq_cl * r * r, 0.0);
  for (auto &tk : rows) {
    int cl = std::atoi(tk[2].c_str());
    double gm = std::atof(tk[3].c_str());
    if (cl > g.nseq_cl) return AwStatus::err(-3, "Fragment_Memory: Wrong sequance separation class tag");
    int is[20], js[20];
    int in = class_indices(g, tk[0], is), jn = class_indices(g, tk[1], js);
    if (in < 0 || jn < 0)
      return AwStatus::err(-3, (in == -2 || jn == -2) ? "Fragment_Memory: Cannot build gamma array" : "Fragment_Memory: Incorrect entery in gamma file");
    for (int i = 0; i < in; ++i)
      for (int j = 0; j < jn; ++j) {
        g.gamma[(size_t)(cl - 1) * r * r + is[i] * r + js[j]] = gm;
        if (is[i] != js[j]) g.gamma[(size_t)(cl - 1) * r * r + js[j] * r + is[i]] = gm;
      }
  }
  return AwStatus();
}

// ---------------------------------------------------------------------------
// .gro files are parsed once and cached; each memory then applies the
// reference's record loop (fm.cpp:61-98) to the cached records.

struct GroRec { int ires; char resty[64], atomty[64]; double x, y, z; };
struct GroFile { bool ok = false; std::vector<GroRec> recs; bool sorted = true; };

static char three_to_one(const char *s)
{
  static const char *nm[20] = {"ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
                               "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL"};
  static const char one[] = "ARNDCQEGHILKMFPSTWYV";
  if (std::strlen(s) == 3)
    for (int i = 0; i < 20; ++i) if (std::strcmp(s, nm[i]) == 0) return one[i];
  return '?';
}

static GroFile load_gro(const std::string &path)
{
  GroFile g;
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return g;
  char buff[201];
  int nAtoms = 0;
  if (!std::fgets(buff, 200, f) || std::fscanf(f, "%d", &nAtoms) != 1) nAtoms = 0;
  if (!std::fgets(buff, 200, f)) nAtoms = 0;
  g.recs.reserve(nAtoms > 0 ? nAtoms : 0);
  for (int i = 0; i < nAtoms; ++i) {
    GroRec r;
    char rt[64], at[64];
    int iatom;
    if (std::fscanf(f, "%d %63s %63s %d %lf %lf %lf", &r.ires, rt, at, &iatom, &r.x, &r.y, &r.z) != 7) break;
    std::snprintf(r.resty, sizeof r.resty, "%s", rt);
    std::snprintf(r.atomty, sizeof r.atomty, "%s", at);
    if (!g.recs.empty() && r.ires < g.recs.back().ires) g.sorted = false;
    g.recs.push_back(r);
  }
  std::fclose(f);
  g.ok = true;
  return g;
}

// returns 0 or Fragment_Memory error code (1 file, 2 atom count)
static int build_memory(const GroFile &g, int pos, int fpos, int len, double weight, AwMemory &m)
{
  if (!g.ok) return 1;
  m.pos = pos; m.len = len; m.weight = weight;
  m.se.assign(len, '\0');
  m.ca.assign((size_t)len * 3, 0.0);
  m.cb.assign((size_t)len * 3, 0.0);
  int nca = 0, ncb = 0;
  size_t k0 = 0;
  if (g.sorted) {  // first record with ires > fpos
    size_t lo = 0, hi = g.recs.size();
    while (lo < hi) { size_t mid = (lo + hi) / 2; if (g.recs[mid].ires > fpos) hi = mid; else lo = mid + 1; }
    k0 = lo;
  }
  for (size_t k = k0; k < g.recs.size(); ++k) {
    const GroRec &r = g.recs[k];
    int ires = r.ires;
    if (ires > fpos + len) break;
    if (ires > fpos && ires <= fpos + len) {
      ires -= fpos + 1;
      double x = r.x * 10, y = r.y * 10, z = r.z * 10;
      if (std::strcmp(r.atomty, "CA") == 0) {
        if (ires >= len || nca >= len) return 2;
        m.se[ires] = three_to_one(r.resty);
        m.ca[3 * ires] = x; m.ca[3 * ires + 1] = y; m.ca[3 * ires + 2] = z;
        nca++;
      }
      if (std::strcmp(r.atomty, "CB") == 0) {
        if (ires >= len || ncb >= len) return 2;
        m.cb[3 * ires] = x; m.cb[3 * ires + 1] = y; m.cb[3 * ires + 2] = z;
        ncb++;
      }
    }
  }
  return nca == len ? 0 : 2;
}

static char *trim(char *s)
{
  while (std::isspace((unsigned char)*s)) s++;
  size_t n = std::strlen(s);
  while (n > 0 && std::isspace((unsigned char)s[n - 1])) s[--n] = '\0';
  return s;
}

// read_mems (fb.cpp:1966-2042): the [Memories] lines of a .mem file, each
// a window of a .gro structure placed at a target position
static AwStatus read_mem_list(const std::string &dir, const std::string &mem_file, const AwSystem &S,
                              std::vector<AwMemory> &out)
{
  FILE *f = std::fopen(join(dir, mem_file).c_str(), "r");
  if (!f) return AwStatus::err(-2, "Fragment_Memory: Error opening mem file");
  std::unordered_map<std::string, GroFile> cache;
  char ln[500];
  int state = 0;  // 0 none, 1 target, 2 memories
  out.clear();
  while (std::fgets(ln, sizeof ln, f)) {
    char *line = trim(ln);
    if (line[0] == '#') continue;
    if (line[0] == '[') state = 0;
    if (blank(line)) { state = 0; continue; }
    if (state == 2) {
      char *str[6];
      int nstr = 0;
      str[nstr] = std::strtok(line, " \t\n");
      while (str[nstr]) {
        nstr++;
        if (nstr > 5) break;
        str[nstr] = std::strtok(nullptr, " \t\n");
      }
      if (nstr != 5) { std::fclose(f); return AwStatus::err(-3, "Fragment_Memory: Error reading mem file"); }
      int tpos = std::atoi(str[1]) - 1, fpos = std::atoi(str[2]) - 1, len = std::atoi(str[3]);
      double w = std::atof(str[4]);
      std::string gname = str[0];
      auto it = cache.find(gname);
      if (it == cache.end()) it = cache.emplace(gname, load_gro(join(dir, gname))).first;
      AwMemory m;
      if (len <= 0 || build_memory(it->second, tpos, fpos, len, w, m)) {
        std::fclose(f);
        return AwStatus::err(-3, "read_mems: Fragment_Memory: Error reading memory fragment " + gname);
      }
      if (tpos < 0 || tpos + len > S.n) {
        std::fclose(f);
        return AwStatus::err(-3, "read_mems: Fragment_Memory: Incorrectly defined memory fragment");
      }
      out.push_back(std::move(m));
    } else if (state == 0) {
      if (std::strcmp(line, "[Target]") == 0) state = 1;
      else if (std::strcmp(line, "[Memories]") == 0) state = 2;
    }
  }
  std::fclose(f);
  return AwStatus();
}

AwStatus aw_read_mems(const std::string &dir, const std::string &mem_file, AwSystem &S)
{
  return read_mem_list(dir, mem_file, S, S.mems);
}

// the fragment frustratometer's decoy library (fb.cpp:794-834): fragments
// to shuffle, or decoy structures; it compares them with the memories,
// which must be loaded
AwStatus aw_read_decoy_mems(const std::string &dir, AwSystem &S)
{
  if (!S.P.ff_flag) return AwStatus();
  if (S.mems.empty())
    return AwStatus::err(-3, "Fragment_Frustratometer: the fragment memories are not loaded (a table read from "
                             "fm_table.energy / fm_table.force reads none)");
  AwStatus s = read_mem_list(dir, S.ff_file, S, S.ff_mems);
  if (!s.ok()) return s;
  if (S.P.ff_mode == 1) S.P.ff_ncalc = (int)S.ff_mems.size() + 1;
  if (S.P.ff_ncalc < 2) return AwStatus::err(-3, "Fragment_Frustratometer: no decoys");
  return AwStatus();
}

AwStatus aw_finish_system(AwSystem &S)
{
  FFParams &P = S.P;
  S.rinfo.assign(S.n, 0);
  for (int c = 0; c < S.nch; ++c)
    for (int k = 0; k < S.ch_len[c]; ++k) {
      int i = S.ch_start[c] + k;
      int ri = aw_restype(S.se[i]) | (c << RI_CHAIN_SHIFT);
      if (k == 0) ri |= RI_FIRST;
      if (k == S.ch_len[c] - 1) ri |= RI_LAST;
      if (S.se[i] == 'G') ri |= RI_GLY;
      if (S.se[i] == 'P') ri |= RI_PRO;
      S.rinfo[i] = ri;
    }
  if (P.phosph_flag)
    for (int j = 0; j < P.n_phosph; ++j) {
      int r = P.phosph_res[j];
      if (r == 0) continue;  // fb.cpp:676: a zero index names no residue
      if (r < 1 || r > S.n) return AwStatus::err(-3, "Phosphorylation: residue index out of range");
      S.rinfo[r - 1] |= RI_PHOS;
    }
  if (P.fm_flag || P.fm_tb_flag) {
    P.fm_min_sep = S.fmg.min_sep();
    P.fm_max_sep = S.fmg.max_sep();
  }
  if (P.fm_flag) {
    for (auto &m : S.mems) {
      int i0 = m.pos, i1 = m.pos + m.len - 1;
      for (int i = i0; i < i1; ++i)
        if (S.rinfo[i] >> RI_CHAIN_SHIFT != S.rinfo[i1] >> RI_CHAIN_SHIFT && i + P.fm_min_sep <= i1)
          return AwStatus::err(-3, "Fragment Memory: Interaction between residues of different chains");
    }
  }
  double dens = 0.0;
  if (P.water_flag) for (int w = 0; w < P.n_wells; ++w) dens = std::fmax(dens, P.w_rmax_th[w]);
  if (P.helix_flag) dens = std::fmax(dens, P.h_rmax_th);
  P.dens_reach = dens;
  double reach = dens;
  if (P.dssp_flag) reach = std::fmax(reach, P.dssp_cut);
  if (P.pap_flag) reach = std::fmax(reach, std::sqrt(P.pap_cutoff_sq));
  double evr = 0.0;
  if (P.ev_on)
    for (int a = 0; a < 6; ++a)
      for (int b = 0; b < 6; ++b)
        if (P.ev_lambda[a][b] != 0.0) evr = std::fmax(evr, std::fmax(P.ev_cut_short[a][b], P.ev_cut_long[a][b]));
  P.ev_reach = evr;
  // the barrier sweeps the residue pairs of the Verlet list: its reach
  if (P.ssb_flag) reach = std::fmax(reach, aw_ssb_cut(P));
  P.pair_reach = std::fmax(reach, evr);
  double lr = P.pair_list_cutoff + P.list_skin;
  P.list_r2 = lr * lr;
  return AwStatus();
}

// ---------------------------------------------------------------------------
// Fragment-memory work description of one system: memories are grouped by
// (pos, len); a group's (i, j) pairs are the ones
// compute_fragment_memory_potential visits (fb.cpp:4110-4125, frag_mem_map
// 784-788) with a non-zero gamma.  rf holds, per (memory, pair), the
// memory distances CA(i)-CA(j), CA(i)-CB(j), CB(i)-CA(j), CB(i)-CB(j)
// (Fragment_Memory::Rf, fm.cpp:190-202); a CB entry of a glycine memory
// residue is stored as -1 (skipped, fb.cpp:4148-4149).

AwStatus aw_build_fm(const AwSystem &S, const FFParams &P, AwFMPack &B)
{
  std::map<std::pair<int, int>, std::vector<int>> by;
  for (size_t k = 0; k < S.mems.size(); ++k) by[{S.mems[k].pos, S.mems[k].len}].push_back((int)k);
  int min_sep = S.fmg.min_sep(), max_sep = S.fmg.max_sep();
  if (S.n >= 32768) return AwStatus::err(-1, "fragment memory supports at most 32767 residues per system");
  for (auto &kv : by) {
    int pos = kv.first.first, len = kv.first.second;
    // candidate pairs of this window (fb.cpp:4110-4125, frag_mem_map 784-788)
    std::vector<std::pair<int, int>> prs;
    for (int di = 0; di < len - min_sep; ++di) {
      int i = pos + di;
      int je = pos + len - 1;
      if (max_sep != -1) je = std::min(je, i + max_sep);
      if (je >= S.n) return AwStatus::err(-3, "Missing residues in memory potential");
      for (int j = i + min_sep; j <= je; ++j) {
        double g = S.fmg.get(aw_restype(S.se[i]), aw_restype(S.se[j]), i, j);
        if (g == 0.0) continue;
        prs.push_back({i - pos, j - pos});
      }
    }
    if (prs.empty()) continue;
    // by separation, then first residue: neighbouring items of a group touch
    // distinct residues, so their force updates do not collide
    std::stable_sort(prs.begin(), prs.end(), [](const std::pair<int, int> &a, const std::pair<int, int> &b) {
      return a.second - a.first < b.second - b.first;
    });
    int np = (int)prs.size(), nmem = (int)kv.second.size();
    long long rf_base = (long long)(B.rf.size() / 4);
    // the kernels address a replica's table with 32-bit byte offsets of
    // 32-byte (FP64) entries
    if (rf_base + (long long)np * nmem >= (1LL << 27))
      return AwStatus::err(-8, "fragment memory table exceeds 2^27 entries");
    if (nmem >= (1 << 24) || np >= (1 << 19)) return AwStatus::err(-1, "too many memories or pairs in one fragment window");
    int coef0 = (int)B.coef.size();
    for (int q = 0; q < np; ++q) {
      int i = pos + prs[q].first, j = pos + prs[q].second;
      double sig2 = std::pow((double)std::abs(i - j), 2 * P.fm_sigma_exp);
      B.items.push_back(S.fmg.get(aw_restype(S.se[i]), aw_restype(S.se[j]), i, j));
      B.items.push_back(1.4426950408889634 / (2.0 * sig2));  // log2(e) / (2 sigma^2)
      B.items.push_back(1.0 / sig2);
      int flags = (S.se[i] == 'G' ? AW_FM_GLY_I : 0) | (S.se[j] == 'G' ? AW_FM_GLY_J : 0);
      int meta[5] = {(i << 16) | j, (int)(rf_base + q), np, nmem | flags, coef0};
      B.meta.insert(B.meta.end(), meta, meta + 5);
    }
    // weights of the group, zero-padded to a multiple of AW_FM_PAD (the
    // kernels read them in aligned blocks)
    for (int mk : kv.second) B.coef.push_back(P.epsilon * P.k_frag_mem * S.mems[mk].weight);
    while (B.coef.size() % AW_FM_PAD) B.coef.push_back(0.0);
    for (int mk : kv.second) {
      const AwMemory &m = S.mems[mk];
      auto R = [](const double *a, const double *b) {
        return std::sqrt((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]));
      };
      for (auto &pr : prs) {
        int a = pr.first, b = pr.second;  // a < b
        const double *cai = &m.ca[3 * a], *caj = &m.ca[3 * b], *cbi = &m.cb[3 * a], *cbj = &m.cb[3 * b];
        bool gi = m.se[a] == 'G', gj = m.se[b] == 'G';
        B.rf.push_back(R(cai, caj));                 // rf[0][a][b], a<b: CA-CA
        B.rf.push_back(gj ? -1.0 : R(cai, cbj));     // rf[1][a][b]: CA(i)-CB(j)
        B.rf.push_back(gi ? -1.0 : R(caj, cbi));     // rf[1][b][a]: CA(j)-CB(i)
        B.rf.push_back((gi || gj) ? -1.0 : R(cbi, cbj));  // rf[0][b][a]: CB-CB
      }
    }
  }
  return AwStatus();
}


// ---------------------------------------------------------------------------
// Fragment-memory table mode.

// compute_fragment_memory_table (fb.cpp:4238-4323) with the allocation of
// fb.cpp:997-1017: for residue i and every memory covering it (pos <= i <
// pos + len - minSep, the frag_mem_map of fb.cpp:775-789), j from i + minSep
// to the memory's end (at most i + maxSep) and the four atom pairs (a CB
// skipped when the target or the memory residue is a glycine), the curve of
// (i, j, k) gets V(r) = -eps k_fm w gamma exp(-(r - rf)^2 / (2 sigma^2)) and
// V (r - rf) / (sigma^2 r) added at every grid point r = rmin + ir dr, with
// sigma^2 = |i - j|^(2 fm_sigma_exp) well_width^2.
AwStatus aw_build_fm_table(AwSystem &S)
{
  const FFParams &P = S.P;
  AwFMTable &T = S.fmt;
  int n = S.n, minsep = S.fmg.min_sep(), maxsep = S.fmg.max_sep();
  T.tb_size = P.tb_size;
  T.tb_nbrs = maxsep != -1 ? maxsep - minsep + 1 : n - minsep;
  if (T.tb_nbrs <= 0 || T.tb_size < 2) return AwStatus::err(-3, "Fragment_Memory_Table: empty table");
  T.cid.assign((size_t)4 * n * T.tb_nbrs, -1);
  T.e.clear();
  T.f.clear();
  int nc = 0;
  for (int i = 0; i < n; ++i) {
    int it = aw_restype(S.se[i]);
    for (const AwMemory &m : S.mems) {
      if (i < m.pos || i >= m.pos + m.len - minsep) continue;
      double ekw = P.epsilon * P.k_frag_mem * m.weight;
      int js = i + minsep, je = m.pos + m.len - 1;
      if (maxsep != -1) je = std::min(je, i + maxsep);
      if (je >= n) return AwStatus::err(-3, "Missing residues in memory potential");
      for (int j = js; j <= je; ++j) {
        double sig2 = std::pow((double)std::abs(i - j), 2 * P.fm_sigma_exp) * P.tb_well_width * P.tb_well_width;
        double ekwg = ekw * S.fmg.get(it, aw_restype(S.se[j]), i, j);
        int a = i - m.pos, b = j - m.pos;
        bool gi = S.se[i] == 'G' || m.se[a] == 'G', gj = S.se[j] == 'G' || m.se[b] == 'G';
        const double *cai = &m.ca[3 * a], *caj = &m.ca[3 * b], *cbi = &m.cb[3 * a], *cbj = &m.cb[3 * b];
        auto R = [](const double *p, const double *q) {
          return std::sqrt((p[0] - q[0]) * (p[0] - q[0]) + (p[1] - q[1]) * (p[1] - q[1]) + (p[2] - q[2]) * (p[2] - q[2]));
        };
        const double rfk[4] = {R(cai, caj), gj ? 0.0 : R(cai, cbj), gi ? 0.0 : R(cbi, caj), (gi || gj) ? 0.0 : R(cbi, cbj)};
        const bool live[4] = {true, !gj, !gi, !gi && !gj};
        for (int k = 0; k < 4; ++k) {
          if (!live[k]) continue;
          size_t itb = 4 * ((size_t)T.tb_nbrs * i + (j - js)) + k;
          if (T.cid[itb] < 0) {
            T.cid[itb] = nc++;
            T.e.resize((size_t)nc * T.tb_size, 0.0);
            T.f.resize((size_t)nc * T.tb_size, 0.0);
          }
          double *ce = &T.e[(size_t)T.cid[itb] * T.tb_size], *cf = &T.f[(size_t)T.cid[itb] * T.tb_size];
          for (int ir = 0; ir < T.tb_size; ++ir) {
            double r = P.tb_rmin + ir * P.tb_dr;
            double dr = r - rfk[k];
            double V = -ekwg * std::exp(-dr * dr / (2 * sig2));
            ce[ir] += V;
            cf[ir] += V * dr / (sig2 * r);
          }
        }
      }
    }
  }
  return AwStatus();
}

// read_fragment_memory_table (fb.cpp:4179-4236): both files hold
// 4 n tb_nbrs rows of tb_size values; every row read becomes a curve (a
// row of zeros, the reference's output for a curve it never allocated, is
// kept as a zero curve: present, contributing nothing)
AwStatus aw_read_fm_table(const std::string &dir, AwSystem &S)
{
  const FFParams &P = S.P;
  AwFMTable &T = S.fmt;
  int n = S.n, minsep = S.fmg.min_sep(), maxsep = S.fmg.max_sep();
  T.tb_size = P.tb_size;
  T.tb_nbrs = maxsep != -1 ? maxsep - minsep + 1 : n - minsep;
  if (T.tb_nbrs <= 0 || T.tb_size < 2) return AwStatus::err(-3, "Fragment_Memory_Table: empty table");
  size_t ntb = (size_t)4 * n * T.tb_nbrs;
  std::vector<double> ev, fv;
  for (int pass = 0; pass < 2; ++pass) {
    std::ifstream in(join(dir, pass ? "fm_table.force" : "fm_table.energy"));
    if (!in) return AwStatus::err(-2, "Fragment memory table files not found!");
    std::vector<double> &v = pass ? fv : ev;
    double val;
    while (in >> val) {
      if (v.size() >= ntb * (size_t)T.tb_size) return AwStatus::err(-3, "Fragment memory table file format error!");
      v.push_back(val);
    }
    if (v.empty()) return AwStatus::err(-3, "Fragment memory table file format error!");
    // the reference's end check (fb.cpp:4210, 4235)
    size_t itb = (v.size() - 1) / T.tb_size, ir = v.size() - itb * T.tb_size;
    if (itb != ntb - 1 && ir != (size_t)T.tb_size) return AwStatus::err(-3, "Fragment memory table file format error!");
  }
  if (fv.size() > ev.size()) return AwStatus::err(-3, "Fragment memory table file format error!");
  fv.resize(ev.size(), 0.0);
  T.cid.assign(ntb, -1);
  T.e.clear();
  T.f.clear();
  size_t rows = (ev.size() + T.tb_size - 1) / T.tb_size;
  int nc = 0;
  for (size_t itb = 0; itb < rows; ++itb) {
    size_t b = itb * T.tb_size, e = std::min(ev.size(), b + T.tb_size);
    bool zero = true;
    for (size_t q = b; q < e; ++q) zero = zero && ev[q] == 0.0 && fv[q] == 0.0;
    if (zero) { T.cid[itb] = -2; continue; }
    T.cid[itb] = nc++;
    T.e.insert(T.e.end(), ev.begin() + b, ev.begin() + e);
    T.f.insert(T.f.end(), fv.begin() + b, fv.begin() + e);
    T.e.resize((size_t)nc * T.tb_size, 0.0);
    T.f.resize((size_t)nc * T.tb_size, 0.0);
  }
  return AwStatus();
}

// output_fragment_memory_table (fb.cpp:6062-6106): "%.12f " per value, one
// row per curve, zeros for a curve never allocated and for an infinity
AwStatus aw_write_fm_table(const std::string &dir, const AwSystem &S)
{
  const AwFMTable &T = S.fmt;
  FILE *fe = std::fopen(join(dir, "fm_table.energy").c_str(), "w");
  FILE *ff = std::fopen(join(dir, "fm_table.force").c_str(), "w");
  if (!fe || !ff) {
    if (fe) std::fclose(fe);
    if (ff) std::fclose(ff);
    return AwStatus::err(-2, "Fragment memory table files not found!");
  }
  for (size_t itb = 0; itb < T.cid.size(); ++itb) {
    int c = T.cid[itb];
    for (int ir = 0; ir < T.tb_size; ++ir) {
      double e = c >= 0 ? T.e[(size_t)c * T.tb_size + ir] : 0.0, f = c >= 0 ? T.f[(size_t)c * T.tb_size + ir] : 0.0;
      if (std::isinf(e)) e = 0.0;
      if (std::isinf(f)) f = 0.0;
      std::fprintf(fe, "%.12f ", e);
      std::fprintf(ff, "%.12f ", f);
    }
    std::fprintf(fe, "\n");
    std::fprintf(ff, "\n");
  }
  std::fclose(fe);
  std::fclose(ff);
  return AwStatus();
}

// the pair path's loop (fb.cpp:7513-7527): j from i + minSep to the last
// residue of i's chain (ch_pos + ch_len - 2 with the reference's 1-based
// chain starts, fb.cpp:546), at most i + maxSep; a pair whose CA-CA curve
// does not exist is skipped (fb.cpp:4341); per atom pair the curve id, -3
// where a target glycine has no CB (fb.cpp:4378-4379)
void aw_fm_table_items(const AwSystem &S, std::vector<int> &items)
{
  items.clear();
  const AwFMTable &T = S.fmt;
  if (T.cid.empty()) return;
  int minsep = S.fmg.min_sep(), maxsep = S.fmg.max_sep();
  for (int c = 0; c < S.nch; ++c) {
    int last = S.ch_start[c] + S.ch_len[c] - 1;
    for (int i = S.ch_start[c]; i <= last; ++i) {
      int jrn = last;
      if (maxsep != -1) jrn = std::min(jrn, i + maxsep);
      for (int j = i + minsep; j <= jrn; ++j) {
        size_t base = 4 * ((size_t)T.tb_nbrs * i + (j - i - minsep));
        if (T.cid[base] == -1) continue;
        bool gi = S.se[i] == 'G', gj = S.se[j] == 'G';
        const bool skip[4] = {false, gj, gi, gi || gj};
        for (int k = 0; k < 4; ++k) {
          if (skip[k]) continue;
          int c = T.cid[base + k];
          const int it[4] = {i, j, k, c >= 0 ? c : -1};
          items.insert(items.end(), it, it + 4);
          if (c == -1) break;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// coeff-toggled terms: side files

// native distance of an AMH-Go pair: the go_rnative*.dat matrices with
// censoring flag 2 (the CA-CB one for both mixed orders, fb.cpp:3904-3909),
// else Fragment_Memory::Rf of amh-go.gro (fm.cpp:190-202)
double aw_go_rnative(const AwSystem &S, int i, int ia, int j, int ja)
{
  const AwExtras &X = S.xt;
  if (S.P.amh_go_censor == 2) {
    size_t k = (size_t)i * S.n + j;
    return X.rn[ia == ja ? ia : 2][k];
  }
  const AwMemory &m = X.amh_mem;
  const double *a, *b;
  if (ia == 0 && ja == 0) { int lo = std::min(i, j), hi = std::max(i, j); a = &m.ca[3 * lo]; b = &m.ca[3 * hi]; }
  else if (ia == 1 && ja == 1) { int lo = std::min(i, j), hi = std::max(i, j); a = &m.cb[3 * hi]; b = &m.cb[3 * lo]; }
  else if (ia == 0) { a = &m.ca[3 * i]; b = &m.cb[3 * j]; }
  else { a = &m.ca[3 * j]; b = &m.cb[3 * i]; }
  if (a == b) return 0.0;
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

AwStatus aw_read_extras(const std::string &dir, const std::string &cr_file, AwSystem &S)
{
  FFParams &P = S.P;
  AwExtras &X = S.xt;
  int n = S.n;
  if (P.cr_flag) {  // read_contact_restraints_file, fb.cpp:1106-1170
    std::ifstream in(join(dir, cr_file));
    if (!in) return AwStatus::err(-2, "File for Contact Restraints potential doesn't exist");
    struct E { int i, j, ord; double w, r0; };
    std::vector<E> es;
    double r0_max = 0.0;
    int a, b;
    double w, r0;
    while (in >> a >> b >> w >> r0) {
      int j1 = std::min(a, b) - 1, j2 = std::max(a, b) - 1;
      if (j1 < 0 || j1 >= n || j2 < 0 || j2 >= n)
        return AwStatus::err(-3, "Indices out of range in Contact Restraints potential file");
      es.push_back(E{j1, j2, (int)es.size(), w * P.k_cont_rest, r0});
      if (r0 > r0_max) r0_max = r0;
    }
    // by lower residue, then upper (fb.cpp:1163-1166; ties in file order)
    std::stable_sort(es.begin(), es.end(), [](const E &x, const E &y) { return x.i != y.i ? x.i < y.i : x.j < y.j; });
    X.cr_i.clear(); X.cr_j.clear(); X.cr_w.clear(); X.cr_r0.clear();
    for (const E &e : es) {
      X.cr_i.push_back(e.i); X.cr_j.push_back(e.j); X.cr_w.push_back(e.w); X.cr_r0.push_back(e.r0);
    }
    P.cr_dr_cutoff = P.cr_sigma * 4.29;
    P.cr_glob_cutoff_sq = std::pow(r0_max + P.cr_dr_cutoff, 2);
    P.pair_list_cutoff = std::fmax(P.pair_list_cutoff, std::sqrt(P.cr_glob_cutoff_sq));
  }
  if (P.memb_flag) {  // fb.cpp:611-620
    std::ifstream in(join(dir, "zim"));
    if (!in) return AwStatus::err(-2, "File zim doesn't exist");
    X.zres.assign(n, 0);
    for (int i = 0; i < n; ++i)
      if (!(in >> X.zres[i])) return AwStatus::err(-3, "Membrane potential parameter file format error");
    // the reference leaves V unset for any other class (fb.cpp:6163-6172)
    for (int i = 0; i < n; ++i)
      if (X.zres[i] < 1 || X.zres[i] > 3) return AwStatus::err(-3, "Membrane potential: zim class outside 1-3");
  }
  if (P.dh_flag) {  // fb.cpp:1018-1046
    std::ifstream in(join(dir, "charge_on_residues.dat"));
    if (!in) return AwStatus::err(-2, "File charge_on_residues.dat doesn't exist");
    X.charge.assign(n, 0.0);
    int total = 0;
    in >> total;
    for (int k = 0; k < total; ++k) {
      int r;
      double q;
      if (!(in >> r >> q)) break;
      if (r < 1 || r > n) return AwStatus::err(-3, "charge_on_residues.dat: residue number out of range");
      X.charge[r - 1] = q;
    }
  }
  if (P.vfm_flag && !(P.fm_flag || P.fm_tb_flag))  // the memories come from those sections (fb.cpp:749-791)
    return AwStatus::err(-3, "Vector_Fragment_Memory needs the memories of [Fragment_Memory]");
  if (P.amhgo_flag) {  // fb.cpp:695-743
    AwStatus s = aw_read_gamma_array(join(dir, "amh-go.gamma"), X.amhg);
    if (!s.ok()) {
      if (s.code == -2) return AwStatus::err(-2, "Cannot read file amh-go.gamma");
      size_t c = s.msg.find(": ");
      return AwStatus::err(s.code, "AMH_Go: " + (c == std::string::npos ? s.msg : s.msg.substr(c + 2)));
    }
    if (X.amhg.frag_resty) return AwStatus::err(-3, "AMH-Go: Wrong call of getGamma() function");
    if (P.amh_go_censor != 2) {
      GroFile g = load_gro(join(dir, "amh-go.gro"));
      int e = build_memory(g, 0, 0, n, 1.0, X.amh_mem);
      if (e == 1) return AwStatus::err(-2, "Cannot read file amh-go.gro");
      if (e == 2) return AwStatus::err(-3, "AMH_Go: Wrong atom count in memory structure file");
    }
    if (P.amh_go_censor == 1) {
      X.censor.assign((size_t)n * n, 0);
      std::ifstream in(join(dir, "frustration_censored_contacts.dat"));
      int a, b;
      while (in >> a >> b) {
        if (a < 1 || a > n || b < 1 || b > n) return AwStatus::err(-3, "AMH-Go: censored contact out of range");
        X.censor[(size_t)(a - 1) * n + (b - 1)] = 1;
      }
    } else if (P.amh_go_censor == 2) {
      const char *names[3] = {"go_rnativeCACA.dat", "go_rnativeCBCB.dat", "go_rnativeCACB.dat"};
      for (int q = 0; q < 3; ++q) {
        std::ifstream in(join(dir, names[q]));
        if (!in) return AwStatus::err(-2, "Go native distance file can't be read");
        X.rn[q].assign((size_t)n * n, 0.0);
        for (size_t k = 0; k < (size_t)n * n; ++k)
          if (!(in >> X.rn[q][k])) return AwStatus::err(-3, "go_rnative*.dat file format error");
      }
    }
    // compute_amhgo_normalization, fb.cpp:3787-3852: after each partner
    // chain, |normi|^p of the running sum (not reset per chain); / 8 N
    double norm = 0.0;
    int iresn = 0;
    for (int ich = 0; ich < S.nch; ++ich) {
      int i0 = S.ch_start[ich];
      iresn = S.ch_start[ich] + S.ch_len[ich];
      for (int i = i0; i < iresn; ++i)
        for (int ia = 0; ia <= (S.se[i] == 'G' ? 0 : 1); ++ia) {
          double normi = 0.0;
          for (int jch = 0; jch < S.nch; ++jch) {
            for (int j = S.ch_start[jch]; j < S.ch_start[jch] + S.ch_len[jch]; ++j)
              for (int ja = 0; ja <= (S.se[j] == 'G' ? 0 : 1); ++ja) {
                if (std::abs(i - j) < X.amhg.min_sep()) continue;
                if (aw_go_rnative(S, i, ia, j, ja) < P.amh_go_rc)
                  normi += X.amhg.get(aw_restype(S.se[i]), aw_restype(S.se[j]), i, j);
              }
            norm += std::pow(std::fabs(normi), P.amh_go_p);
          }
        }
    }
    X.go_norm = norm / (8 * iresn);
  }
  return AwStatus();
}

// Device items.  AMH-Go: for every atom a = 2 i + k (CA; CB unless glycine)
// the directed pairs a -> b the reference's force loop can count
// (fb.cpp:3895-3950): b a CA or non-glycine CB, not censored, |i - j| >=
// minSep or another chain, native distance < rc, nonzero gamma (a zero one
// adds nothing).  Contact restraints: the first entry of each residue pair
// (cr_contact_search, fb.cpp:8118-8135) whose sites are a density pair
// (another chain or |i - j| > 1, fb.cpp:7985).  Debye-Hueckel: residue pairs
// i < j with |i - j| > 1 (fb.cpp:8116) and >= the minimum separation
// (6309) whose charge product is nonzero, k by the signs (6334-6343).
// Vector fragment memory: per memory covering i (frag_mem_map), j from i +
// minSep to the memory's end (at most i + maxSep), no glycine in the target
// or the memory (fb.cpp:4008-4027), with the memory's angle
// (Fragment_Memory::VM, fm.cpp:166, 239-243).
AwStatus aw_build_extras(const AwSystem &S, const FFParams &P, AwXtPack &X)
{
  const AwExtras &E = S.xt;
  int n = S.n;
  // the force field's terms need the system's side data (a server's force
  // field with a job context read from another coeff file)
  if ((P.amhgo_flag && !S.P.amhgo_flag) || (P.cr_flag && !S.P.cr_flag) || (P.dh_flag && !S.P.dh_flag) ||
      (P.memb_flag && !S.P.memb_flag))
    return AwStatus::err(-1, "the system lacks the side files of the force field's optional terms");
  auto chain = [&](int i) { return S.rinfo[i] >> RI_CHAIN_SHIFT; };
  X = AwXtPack();
  if (P.amhgo_flag) {
    X.go_start.assign(2 * n + 1, 0);
    int minsep = E.amhg.min_sep();
    for (int a = 0; a < 2 * n; ++a) {
      int i = a >> 1, ka = a & 1;
      X.go_start[a] = (int)X.go.size();
      if (ka == 1 && S.se[i] == 'G') continue;
      int it = aw_restype(S.se[i]);
      for (int j = 0; j < n; ++j) {
        if (!E.censor.empty() && (E.censor[(size_t)i * n + j] || E.censor[(size_t)j * n + i])) continue;
        if (!(std::abs(i - j) >= minsep || chain(i) != chain(j))) continue;
        for (int kb = 0; kb < 2; ++kb) {
          if (kb == 1 && S.se[j] == 'G') continue;
          if (j == i && kb == ka) continue;
          double rn = aw_go_rnative(S, i, ka, j, kb);
          if (!(rn < P.amh_go_rc)) continue;
          double g = E.amhg.get(it, aw_restype(S.se[j]), i, j);
          if (g == 0.0) continue;
          XtGo q;
          q.b = 2 * j + kb; q.pad = 0;
          q.rn = rn; q.gamma = g; q.s2 = std::pow(std::abs(i - j), 0.3);
          X.go.push_back(q);
        }
      }
    }
    X.go_start[2 * n] = (int)X.go.size();
  }
  if (P.cr_flag) {
    for (size_t k = 0; k < E.cr_i.size(); ++k) {
      if (k > 0 && E.cr_i[k] == E.cr_i[k - 1] && E.cr_j[k] == E.cr_j[k - 1]) continue;  // not the first
      int i = E.cr_i[k], j = E.cr_j[k];
      if (!(chain(i) != chain(j) || j - i > 1)) continue;
      X.cr.push_back(XtPair{i, j, E.cr_w[k], E.cr_r0[k]});
    }
  }
  if (P.dh_flag) {
    for (int i = 0; i < n; ++i) {
      double qi = E.charge[i];
      if (qi == 0.0) continue;
      for (int j = i + 2; j < n; ++j) {
        double qj = E.charge[j];
        if (qj == 0.0 || j - i < P.dh_min_sep) continue;
        double k = (qi > 0 && qj > 0) ? P.dh_k[0] : ((qi < 0 && qj < 0) ? P.dh_k[1] : P.dh_k[2]);
        X.dh.push_back(XtPair{i, j, k * qi * qj, 0.0});
      }
    }
  }
  if (P.vfm_flag) {
    int minsep = S.fmg.min_sep(), maxsep = S.fmg.max_sep();
    for (const AwMemory &m : S.mems) {
      for (int i = m.pos; i < m.pos + m.len - minsep; ++i) {
        int je = m.pos + m.len - 1;
        if (maxsep != -1) je = std::min(je, i + maxsep);
        if (je >= n) continue;  // the reference aborts ("Missing residues in memory potential")
        for (int j = i + minsep; j <= je; ++j) {
          int mi = i - m.pos, mj = j - m.pos;
          if (S.se[i] == 'G' || S.se[j] == 'G' || m.se[mi] == 'G' || m.se[mj] == 'G') continue;
          const double *a1 = &m.ca[3 * mi], *b1 = &m.cb[3 * mi], *a2 = &m.ca[3 * mj], *b2 = &m.cb[3 * mj];
          double u[3] = {b1[0] - a1[0], b1[1] - a1[1], b1[2] - a1[2]}, v[3] = {b2[0] - a2[0], b2[1] - a2[1], b2[2] - a2[2]};
          double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
          double nu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
          X.vfm.push_back(XtPair{i, j, std::acos(uv / (nu * nv)), 0.0});
        }
      }
    }
  }
  return AwStatus();
}

// the coeff-toggled terms' parameters two systems of one batch must share
// (the kernels read them from the batch's single parameter block); the
// per-system values (phosphorylated residues, the contact restraints'
// global cutoff, the list reach) travel with each replica instead
bool aw_same_xt_params(const FFParams &a, const FFParams &b)
{
  auto same = [](const void *x, const void *y, size_t n) { return std::memcmp(x, y, n) == 0; };
#define AW_SAME(f) same(&a.f, &b.f, sizeof a.f)
  return AW_SAME(amhgo_flag) && AW_SAME(cr_flag) && AW_SAME(vfm_flag) && AW_SAME(ssb_flag) && AW_SAME(memb_flag) &&
         AW_SAME(phosph_flag) && AW_SAME(dh_flag) && AW_SAME(k_amh_go) && AW_SAME(amh_go_p) && AW_SAME(amh_go_rc) &&
         AW_SAME(k_cont_rest) && AW_SAME(cr_sigma_sq_inv) && AW_SAME(cr_dr_cutoff) && AW_SAME(k_vfm) &&
         AW_SAME(vfm_sigma_sq) && AW_SAME(k_ssb1) && AW_SAME(ssb_rmin1) && AW_SAME(ssb_rmax1) && AW_SAME(k_ssb2) &&
         AW_SAME(ssb_rmin2) && AW_SAME(ssb_rmax2) && AW_SAME(ssb_kappa) && AW_SAME(ssb_rshift) &&
         AW_SAME(ssb_ij_sep) && AW_SAME(ssb_rad_cor) && AW_SAME(k_memb) && AW_SAME(memb_kbin) && AW_SAME(memb_xo) &&
         AW_SAME(memb_len) && AW_SAME(memb_rho0_max) && AW_SAME(memb_rho0_distor) && AW_SAME(g_memb) &&
         AW_SAME(dh_k) && AW_SAME(dh_kscreen) && AW_SAME(dh_length) && AW_SAME(phosph_gamma) && AW_SAME(epsilon);
#undef AW_SAME
}
