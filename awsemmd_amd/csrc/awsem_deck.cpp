// awsem_deck.cpp -- LAMMPS input decks for the FixBackbone path.
//
// The shipped AWSEM decks (examples/examples_new/*/folding.in and the like)
// drive `fix backbone` from a LAMMPS input script: `read_data` of an
// `atom_style awsemmd` data file (src/atom_vec_awsemmd.cpp:56: id, molecule,
// residue, type, q, x y z), groups for the three atom kinds, the
// excluded-volume pair style (src/pair_excluded_volume.cpp:175-265), harmonic
// bonds, `velocity create`, `fix nve` + `fix langevin`, `fix backbone`
// (fb.cpp:86-106: the group, beta and oxygen groups, coeff and sequence
// files), thermo output, `dump atom`, and `fix print` of equal-style
// variables built from f_ID[k] (compute_vector, fb.cpp:8223-8232), emol,
// epair and ke.  This file interprets that subset on top of the library's
// own C ABI: the force field is an awsem context, the dynamics one replica
// of an awsem batch (which continues exactly from its stored forces across
// the output intervals), and FixBackbone's energy.log is written at the
// steps the reference writes it (fb.cpp:7661-7672).  A command outside the
// subset is refused with AWSEM_E_UNSUPPORTED naming it and its line.
//
// LAMMPS itself is not part of the reference; what LAMMPS does around the
// fix (the random streams of `velocity create` and `fix langevin`, thermo
// and dump number formats, the shrink-wrapped box) follows its documented
// behaviour with this library's own generators and formats.

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/awsem.h"

namespace {

// LAMMPS real units (update.cpp force->boltz, mvv2e, ftm2v)
const double kBoltz = 0.0019872067;
const double kMvv2e = 48.88821291 * 48.88821291;

struct DeckErr {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg) { throw DeckErr{code, msg}; }

struct Atom {
  long long id;
  int mol, res, type;
  double q;
  double x[3], v[3];
};

struct Tok {
  std::string s;
  bool quoted;
};

struct Var {
  std::string style;  // equal, index, string
  std::string text;
};

struct Dump {
  std::string id, group, file;
  long long every = 0;
  long long last = -1;  // step of the last frame (a run's setup repeats no frame)
  bool sort_id = false;
  FILE *fp = nullptr;
};

struct PrintFix {
  std::string id, text, file, title;
  long long every = 0;
  bool screen = true;
  FILE *fp = nullptr;
};

uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// uniform in [0, 1) from (seed, atom id, draw)
double urand3(uint64_t seed, uint64_t id, uint64_t k)
{
  uint64_t h = mix64(mix64(mix64(seed) ^ id) ^ k);
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

double to_d(const std::string &s, const std::string &what)
{
  char *end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  if (s.empty() || *end) fail(AWSEM_E_FORMAT, "Expected a floating point value for " + what + ": '" + s + "'");
  return v;
}

long long to_ll(const std::string &s, const std::string &what)
{
  char *end = nullptr;
  long long v = std::strtoll(s.c_str(), &end, 10);
  if (s.empty() || *end) fail(AWSEM_E_FORMAT, "Expected an integer for " + what + ": '" + s + "'");
  return v;
}

// LAMMPS type bounds ("3", "*", "2*", "*4", "2*4"), utils::bounds
void bounds(const std::string &s, int nmax, int &lo, int &hi)
{
  size_t p = s.find('*');
  if (p == std::string::npos) {
    lo = hi = (int)to_ll(s, "a type");
  } else {
    lo = p == 0 ? 1 : (int)to_ll(s.substr(0, p), "a type range");
    hi = p + 1 >= s.size() ? nmax : (int)to_ll(s.substr(p + 1), "a type range");
  }
  if (lo < 1 || hi > nmax || lo > hi) fail(AWSEM_E_FORMAT, "Numeric index " + s + " is out of bounds");
}

std::string fmt15g(double v)
{
  char b[64];
  std::snprintf(b, sizeof b, "%.15g", v);
  return b;
}

}  // namespace

struct awsem_deck {
  std::string path, dir, err;
  bool broken = false;             // the script failed while being opened
  awsem_deck_opts opts{};
  std::vector<std::string> lines;  // logical lines (continuations joined)
  std::vector<int> line_no;
  size_t pc = 0;                   // next line to execute

  // settings
  double dt = 1.0;                 // fs (LAMMPS real units default)
  double neigh_skin = 2.0;         // A: `neighbor skin style` (LAMMPS real units default)
  char boundary[3] = {'p', 'p', 'p'};
  bool units_real = false, style_awsemmd = false, bond_harmonic = false, special_fene = false;
  bool pair_on = false;
  int pair_p = 2;
  double pair_cs = 0.0, pair_cl = 0.0;
  std::vector<double> lam, pcs, pcl;  // [(ntypes+1)^2]
  std::vector<int> pset;
  long long step = 0;
  long long thermo_every = 0;
  std::vector<std::string> thermo_keys{"step", "temp", "epair", "emol", "etotal"};

  // system
  bool have_data = false;
  int ntypes = 0, nbtypes = 0;
  double box[3][2] = {{-0.5, 0.5}, {-0.5, 0.5}, {-0.5, 0.5}};
  std::vector<double> mass;        // [ntypes + 1]
  std::vector<Atom> atoms;         // file order
  std::map<long long, int> index_of;
  std::vector<double> bond_k, bond_r0;  // [nbtypes + 1]
  std::vector<int> bond_type, bond_a, bond_b;  // atom ids
  std::map<std::string, std::vector<char>> groups;  // per atom (file order)

  // fixes
  std::string fb_id, fb_group, fb_beta, fb_oxy, fb_coeff, fb_seq;
  bool have_nve = false, have_langevin = false;
  long long tf_every = 0, nf_every = 0, ff_every = 0;  // the frustratometers' output intervals (0: none)
  double lg_t0 = 0.0, lg_t1 = 0.0, lg_damp = 0.0;
  long long lg_seed = 0;
  std::vector<Dump> dumps;
  std::vector<PrintFix> prints;
  std::map<std::string, Var> vars;

  // engine
  awsem_ctx *ctx = nullptr;
  awsem_batch *batch = nullptr;
  int n_res = 0;
  std::vector<int> atom_of;        // [3][n_res] -> atom index (file order)
  double energy[AWSEM_N_ET] = {0};
  bool energy_valid = false;
  bool state_dirty = true;         // host x/v newer than the batch's
  FILE *thermo_fp = nullptr;
  bool elog_open = false;

  ~awsem_deck()
  {
    for (Dump &d : dumps) if (d.fp) std::fclose(d.fp);
    for (PrintFix &p : prints) if (p.fp) std::fclose(p.fp);
    if (thermo_fp) std::fclose(thermo_fp);
    if (batch) awsem_batch_destroy(batch);
    if (ctx) awsem_destroy(ctx);
  }

  std::string here() const
  {
    int ln = pc > 0 && pc - 1 < line_no.size() ? line_no[pc - 1] : 0;
    return " (" + path + ":" + std::to_string(ln) + ")";
  }

  std::string resolve(const std::string &f) const { return !f.empty() && f[0] == '/' ? f : dir + "/" + f; }

  // ---- reading the script ---------------------------------------------

  void load(const std::string &p)
  {
    path = p;
    size_t s = p.find_last_of('/');
    dir = s == std::string::npos ? "." : p.substr(0, s);
    std::ifstream in(p);
    if (!in) fail(AWSEM_E_FILE, "Cannot open input script " + p);
    std::string line, acc;
    int no = 0, start = 0;
    while (std::getline(in, line)) {
      ++no;
      if (acc.empty()) start = no;
      // a trailing '&' continues the command on the next line
      size_t e = line.find_last_not_of(" \t\r");
      if (e != std::string::npos && line[e] == '&') {
        acc += line.substr(0, e) + " ";
        continue;
      }
      acc += line;
      lines.push_back(acc);
      line_no.push_back(start);
      acc.clear();
    }
    if (!acc.empty()) { lines.push_back(acc); line_no.push_back(start); }
  }

  // comments end a line outside quotes; ${name} and $x are replaced outside
  // quotes (variables in quoted text, as in fix print, are replaced when
  // the text is printed); quotes group words into one argument
  std::vector<Tok> tokenize(const std::string &raw)
  {
    std::vector<Tok> out;
    size_t i = 0, n = raw.size();
    std::string cur;
    bool in_tok = false, quoted = false;
    auto flush = [&]() {
      if (in_tok) out.push_back({cur, quoted});
      cur.clear();
      in_tok = false;
      quoted = false;
    };
    while (i < n) {
      char c = raw[i];
      if (c == '#') break;
      if (c == '"' || c == '\'') {
        size_t j = raw.find(c, i + 1);
        if (j == std::string::npos) fail(AWSEM_E_FORMAT, "Unbalanced quotes in input line" + here());
        cur += raw.substr(i + 1, j - i - 1);
        in_tok = true;
        quoted = true;
        i = j + 1;
        continue;
      }
      if (c == '$') {
        std::string name;
        if (i + 1 < n && raw[i + 1] == '{') {
          size_t j = raw.find('}', i + 2);
          if (j == std::string::npos) fail(AWSEM_E_FORMAT, "Invalid variable name in input line" + here());
          name = raw.substr(i + 2, j - i - 2);
          i = j + 1;
        } else if (i + 1 < n) {
          name = raw.substr(i + 1, 1);
          i += 2;
        } else {
          fail(AWSEM_E_FORMAT, "Invalid variable name in input line" + here());
        }
        cur += var_text(name);
        in_tok = true;
        continue;
      }
      if (std::isspace((unsigned char)c)) {
        flush();
        ++i;
        continue;
      }
      cur += c;
      in_tok = true;
      ++i;
    }
    flush();
    return out;
  }

  std::string var_text(const std::string &name)
  {
    auto it = vars.find(name);
    if (it == vars.end()) fail(AWSEM_E_FORMAT, "Substitution for illegal variable " + name + here());
    if (it->second.style == "equal") return fmt15g(eval(it->second.text));
    return it->second.text;
  }

  // substitution inside quoted text at output time
  std::string substitute(const std::string &t)
  {
    std::string out;
    for (size_t i = 0; i < t.size();) {
      if (t[i] == '$' && i + 1 < t.size()) {
        std::string name;
        if (t[i + 1] == '{') {
          size_t j = t.find('}', i + 2);
          if (j == std::string::npos) fail(AWSEM_E_FORMAT, "Invalid variable name in " + t);
          name = t.substr(i + 2, j - i - 2);
          i = j + 1;
        } else {
          name = t.substr(i + 1, 1);
          i += 2;
        }
        out += var_text(name);
      } else {
        out += t[i++];
      }
    }
    return out;
  }

  // ---- equal-style expressions: numbers, + - * / ^, parentheses, unary
  // minus, v_name, f_ID and f_ID[k] of fix backbone, and the thermo
  // keywords step, elapsed, dt, time, atoms, temp, ke, pe, etotal, epair,
  // evdwl, emol, ebond
  struct Expr {
    awsem_deck *d;
    const std::string &s;
    size_t i = 0;
    int depth;
    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
    double expr()
    {
      double v = term();
      for (;;) {
        ws();
        if (i < s.size() && (s[i] == '+' || s[i] == '-')) {
          char op = s[i++];
          double r = term();
          v = op == '+' ? v + r : v - r;
        } else {
          return v;
        }
      }
    }
    double term()
    {
      double v = power();
      for (;;) {
        ws();
        if (i < s.size() && (s[i] == '*' || s[i] == '/')) {
          char op = s[i++];
          double r = power();
          if (op == '/' && r == 0.0) fail(AWSEM_E_FORMAT, "Divide by 0 in variable formula " + s);
          v = op == '*' ? v * r : v / r;
        } else {
          return v;
        }
      }
    }
    double power()
    {
      double v = unary();
      ws();
      if (i < s.size() && s[i] == '^') {
        ++i;
        v = std::pow(v, power());
      }
      return v;
    }
    double unary()
    {
      ws();
      if (i < s.size() && s[i] == '-') { ++i; return -unary(); }
      if (i < s.size() && s[i] == '+') { ++i; return unary(); }
      return atom();
    }
    double atom()
    {
      ws();
      if (i >= s.size()) fail(AWSEM_E_FORMAT, "Invalid syntax in variable formula " + s);
      if (s[i] == '(') {
        ++i;
        double v = expr();
        ws();
        if (i >= s.size() || s[i] != ')') fail(AWSEM_E_FORMAT, "Invalid syntax in variable formula " + s);
        ++i;
        return v;
      }
      if (std::isdigit((unsigned char)s[i]) || s[i] == '.') {
        size_t j = i;
        while (j < s.size() && (std::isdigit((unsigned char)s[j]) || s[j] == '.')) ++j;
        if (j < s.size() && (s[j] == 'e' || s[j] == 'E')) {
          size_t k = j + 1;
          if (k < s.size() && (s[k] == '+' || s[k] == '-')) ++k;
          if (k < s.size() && std::isdigit((unsigned char)s[k])) {
            j = k;
            while (j < s.size() && std::isdigit((unsigned char)s[j])) ++j;
          }
        }
        double v = to_d(s.substr(i, j - i), "a number in " + s);
        i = j;
        return v;
      }
      size_t j = i;
      while (j < s.size() && (std::isalnum((unsigned char)s[j]) || s[j] == '_')) ++j;
      std::string name = s.substr(i, j - i);
      i = j;
      if (name.empty()) fail(AWSEM_E_FORMAT, "Invalid syntax in variable formula " + s);
      int idx = -1;
      ws();
      if (i < s.size() && s[i] == '[') {
        size_t k = s.find(']', i);
        if (k == std::string::npos) fail(AWSEM_E_FORMAT, "Invalid syntax in variable formula " + s);
        idx = (int)to_ll(s.substr(i + 1, k - i - 1), "an index in " + s);
        i = k + 1;
      }
      return d->value(name, idx, depth);
    }
  };

  double eval(const std::string &text, int depth = 0)
  {
    if (depth > 32) fail(AWSEM_E_FORMAT, "Variable formulas nest too deeply: " + text);
    Expr e{this, text, 0, depth};
    double v = e.expr();
    e.ws();
    if (e.i != text.size()) fail(AWSEM_E_FORMAT, "Invalid syntax in variable formula " + text);
    return v;
  }

  double fb_total() const
  {
    double t = 0.0;
    for (int k = 1; k < 18; ++k) t += k == AWSEM_ET_VEXCLUDED ? 0.0 : energy[k];
    return t;
  }

  double kinetic() const
  {
    double s = 0.0;
    for (const Atom &a : atoms) s += mass[a.type] * (a.v[0] * a.v[0] + a.v[1] * a.v[1] + a.v[2] * a.v[2]);
    return 0.5 * kMvv2e * s;
  }

  double temperature() const
  {
    double dof = 3.0 * (double)atoms.size() - 3.0;
    return dof > 0 ? 2.0 * kinetic() / (dof * kBoltz) : 0.0;
  }

  double epair() const { return pair_on ? energy[AWSEM_ET_VEXCLUDED] : 0.0; }
  double emol() const { return energy[AWSEM_ET_BOND]; }
  double pe() const { return fb_total() + epair() + emol(); }

  double value(const std::string &name, int idx, int depth)
  {
    if (name.compare(0, 2, "v_") == 0) {
      auto it = vars.find(name.substr(2));
      if (it == vars.end()) fail(AWSEM_E_FORMAT, "Invalid variable name in variable formula: " + name);
      if (it->second.style == "equal") return eval(it->second.text, depth + 1);
      return to_d(it->second.text, "variable " + name);
    }
    if (name.compare(0, 2, "f_") == 0) {
      if (name.substr(2) != fb_id) fail(AWSEM_E_UNSUPPORTED, "Only fix backbone values can be used in formulas: " + name);
      if (idx < 0) return fb_total();   // compute_scalar, fb.cpp:8208-8219
      if (idx < 1 || idx > 17) fail(AWSEM_E_FORMAT, "Variable formula fix vector is accessed out-of-range: " + name);
      return idx == AWSEM_ET_VEXCLUDED ? 0.0 : energy[idx];  // compute_vector(idx - 1), fb.cpp:8223-8232
    }
    if (idx >= 0) fail(AWSEM_E_UNSUPPORTED, "Indexed value not supported in formulas: " + name);
    if (name == "step") return (double)step;
    if (name == "dt") return dt;
    if (name == "time") return dt * (double)step;
    if (name == "atoms") return (double)atoms.size();
    if (name == "temp") return temperature();
    if (name == "ke") return kinetic();
    if (name == "pe") return pe();
    if (name == "etotal") return pe() + kinetic();
    if (name == "epair" || name == "evdwl") return epair();
    if (name == "emol" || name == "ebond") return emol();
    fail(AWSEM_E_UNSUPPORTED, "Keyword " + name + " in variable formula is not supported by this front end");
  }

  // ---- commands ----------------------------------------------------------

  void need(const std::vector<Tok> &a, size_t n, const std::string &cmd)
  {
    if (a.size() < n) fail(AWSEM_E_FORMAT, "Illegal " + cmd + " command" + here());
  }

  [[noreturn]] void unsupported(const std::string &what)
  {
    fail(AWSEM_E_UNSUPPORTED, what + " is not supported by this front end" + here());
  }

  // executes lines until `run` (not executed) or the end; true at a run
  bool exec_until_run()
  {
    while (pc < lines.size()) {
      std::vector<Tok> a = tokenize(lines[pc]);
      if (a.empty()) { ++pc; continue; }
      if (a[0].s == "run") return true;
      ++pc;
      command(a);
    }
    return false;
  }

  void command(const std::vector<Tok> &a)
  {
    const std::string &c = a[0].s;
    if (c == "units") {
      need(a, 2, c);
      if (a[1].s != "real") unsupported("units " + a[1].s + " (only real)");
      units_real = true;
    } else if (c == "dimension") {
      need(a, 2, c);
      if (a[1].s != "3") unsupported("dimension " + a[1].s);
    } else if (c == "timestep") {
      need(a, 2, c);
      dt = to_d(a[1].s, "timestep");
    } else if (c == "boundary") {
      need(a, 4, c);
      for (int d = 0; d < 3; ++d) {
        const std::string &b = a[1 + d].s;
        if (b.size() != 1 && !(b.size() == 2 && b[0] == b[1])) unsupported("boundary " + b);
        if (b[0] != 'p' && b[0] != 's' && b[0] != 'f' && b[0] != 'm') fail(AWSEM_E_FORMAT, "Illegal boundary command" + here());
        boundary[d] = b[0];
      }
    } else if (c == "neighbor") {
      // the engine keeps its own lists; the skin sets how far the terms
      // that take their pairs from FixBackbone's list reach (a list just
      // built holds pair_list_cutoff + skin)
      need(a, 3, c);
      neigh_skin = to_d(a[1].s, c);
      if (neigh_skin < 0.0) fail(AWSEM_E_FORMAT, "Illegal neighbor command" + here());
    } else if (c == "neigh_modify" || c == "atom_modify" || c == "processors" || c == "newton" || c == "comm_modify") {
      // neighbour lists, sorting and decomposition are the engine's own
    } else if (c == "special_bonds") {
      // the excluded-volume pairs leave out exactly the bonded (1-2) pairs
      bool ok = a.size() == 2 && a[1].s == "fene";
      if (a.size() == 5 && (a[1].s == "lj/coul" || a[1].s == "lj") && to_d(a[2].s, c) == 0.0 &&
          to_d(a[3].s, c) == 1.0 && to_d(a[4].s, c) == 1.0)
        ok = true;
      if (!ok) unsupported("special_bonds other than fene (1-2 pairs excluded, 1-3 and 1-4 kept)");
      special_fene = true;
    } else if (c == "atom_style") {
      need(a, 2, c);
      if (a[1].s != "awsemmd") unsupported("atom_style " + a[1].s);
      style_awsemmd = true;
    } else if (c == "bond_style") {
      need(a, 2, c);
      if (a[1].s != "harmonic") unsupported("bond_style " + a[1].s);
      bond_harmonic = true;
    } else if (c == "pair_style") {
      need(a, 2, c);
      if (a[1].s != "vexcluded") unsupported("pair_style " + a[1].s);
      // PairExcludedVolume::settings, pev.cpp:175-195
      if (a.size() != 5) fail(AWSEM_E_FORMAT, "Illegal pair_style command" + here());
      pair_on = true;
      pair_p = (int)to_ll(a[2].s, "pair_style p");
      pair_cs = to_d(a[3].s, "pair_style cutoff");
      pair_cl = to_d(a[4].s, "pair_style cutoff");
    } else if (c == "read_data") {
      need(a, 2, c);
      read_data(resolve(a[1].s));
    } else if (c == "pair_coeff") {
      pair_coeff(a);
    } else if (c == "bond_coeff") {
      need(a, 4, c);
      if (!have_data) fail(AWSEM_E_STATE, "Bond_coeff command before simulation box is defined" + here());
      int lo, hi;
      bounds(a[1].s, nbtypes, lo, hi);
      for (int t = lo; t <= hi; ++t) { bond_k[t] = to_d(a[2].s, c); bond_r0[t] = to_d(a[3].s, c); }
    } else if (c == "mass") {
      need(a, 3, c);
      if (!have_data) fail(AWSEM_E_STATE, "Mass command before simulation box is defined" + here());
      int lo, hi;
      bounds(a[1].s, ntypes, lo, hi);
      for (int t = lo; t <= hi; ++t) mass[t] = to_d(a[2].s, c);
    } else if (c == "velocity") {
      velocity(a);
    } else if (c == "group") {
      group(a);
    } else if (c == "fix") {
      fix(a);
    } else if (c == "unfix") {
      need(a, 2, c);
      unfix(a[1].s);
    } else if (c == "thermo") {
      need(a, 2, c);
      thermo_every = to_ll(a[1].s, "thermo");
    } else if (c == "thermo_style") {
      need(a, 2, c);
      if (a[1].s == "one") {
        thermo_keys = {"step", "temp", "epair", "emol", "etotal"};
      } else if (a[1].s == "custom") {
        thermo_keys.clear();
        for (size_t k = 2; k < a.size(); ++k) {
          static const std::set<std::string> ok = {"step", "elapsed", "dt", "time", "atoms", "temp", "ke",
                                                   "pe", "etotal", "epair", "evdwl", "emol", "ebond"};
          const std::string &w = a[k].s;
          if (!ok.count(w) && w.compare(0, 2, "v_") != 0 && w.compare(0, 2, "f_") != 0)
            unsupported("thermo keyword " + w);
          thermo_keys.push_back(w);
        }
      } else {
        unsupported("thermo_style " + a[1].s);
      }
    } else if (c == "thermo_modify") {
      // number formats and flushing only
    } else if (c == "dump") {
      need(a, 6, c);
      if (a[3].s != "atom") unsupported("dump style " + a[3].s);
      Dump d;
      d.id = a[1].s;
      d.group = a[2].s;
      d.every = to_ll(a[4].s, "dump frequency");
      d.file = resolve(a[5].s);
      if (d.every <= 0) fail(AWSEM_E_FORMAT, "Illegal dump command" + here());
      if (!groups.count(d.group) && d.group != "all") fail(AWSEM_E_FORMAT, "Could not find dump group ID " + d.group + here());
      dumps.push_back(d);
    } else if (c == "dump_modify") {
      need(a, 2, c);
      Dump *d = nullptr;
      for (Dump &x : dumps) if (x.id == a[1].s) d = &x;
      if (!d) fail(AWSEM_E_FORMAT, "Could not find dump_modify ID " + a[1].s + here());
      for (size_t k = 2; k + 1 < a.size(); k += 2) {
        if (a[k].s == "sort") d->sort_id = a[k + 1].s == "id";
        else unsupported("dump_modify " + a[k].s);
      }
    } else if (c == "variable") {
      need(a, 3, c);
      const std::string &st = a[2].s;
      if (st == "equal" || st == "string" || st == "index") {
        need(a, 4, c);
        std::string text;
        for (size_t k = 3; k < a.size(); ++k) text += (k > 3 ? " " : "") + a[k].s;
        if (st == "index") text = a[3].s;  // the first value of the list
        vars[a[1].s] = Var{st == "index" ? "index" : st, text};
      } else if (st == "delete") {
        vars.erase(a[1].s);
      } else {
        unsupported("variable style " + st);
      }
    } else if (c == "reset_timestep") {
      need(a, 2, c);
      step = to_ll(a[1].s, "reset_timestep");
    } else if (c == "print") {
      need(a, 2, c);
      if (!opts.quiet) std::printf("%s\n", substitute(a[1].s).c_str());
    } else {
      unsupported("command " + c);
    }
  }

  void read_data(const std::string &file)
  {
    if (!style_awsemmd) fail(AWSEM_E_STATE, "read_data before atom_style awsemmd" + here());
    std::ifstream in(file);
    if (!in) fail(AWSEM_E_FILE, "Cannot open file " + file + here());
    std::string line;
    std::getline(in, line);  // title
    long long natoms = 0, nbonds = 0;
    // header: counts and box, until the first section keyword
    std::string section;
    while (std::getline(in, line)) {
      std::string t = line.substr(0, line.find('#'));
      std::istringstream ls(t);
      std::vector<std::string> w;
      for (std::string s; ls >> s;) w.push_back(s);
      if (w.empty()) continue;
      if (!std::isdigit((unsigned char)w[0][0]) && w[0][0] != '-' && w[0][0] != '+' && w[0][0] != '.') {
        section = t.substr(t.find_first_not_of(" \t"));
        section = section.substr(0, section.find_last_not_of(" \t\r") + 1);
        break;
      }
      if (w.size() >= 2 && w[1] == "atoms") natoms = to_ll(w[0], "atoms");
      else if (w.size() >= 2 && w[1] == "bonds") nbonds = to_ll(w[0], "bonds");
      else if (w.size() >= 3 && w[1] == "atom" && w[2] == "types") ntypes = (int)to_ll(w[0], "atom types");
      else if (w.size() >= 3 && w[1] == "bond" && w[2] == "types") nbtypes = (int)to_ll(w[0], "bond types");
      else if (w.size() >= 2 && (w[1] == "angles" || w[1] == "dihedrals" || w[1] == "impropers")) {
        if (to_ll(w[0], w[1]) != 0) unsupported("data file " + w[1]);
      } else if (w.size() >= 3 && (w[2] == "types")) {
        // angle/dihedral/improper types: allowed when there are none of them
      } else if (w.size() >= 4 && w[2] == "xlo") { box[0][0] = to_d(w[0], "xlo"); box[0][1] = to_d(w[1], "xhi"); }
      else if (w.size() >= 4 && w[2] == "ylo") { box[1][0] = to_d(w[0], "ylo"); box[1][1] = to_d(w[1], "yhi"); }
      else if (w.size() >= 4 && w[2] == "zlo") { box[2][0] = to_d(w[0], "zlo"); box[2][1] = to_d(w[1], "zhi"); }
      else if (w.size() >= 4 && w[3] == "xy") unsupported("triclinic box");
      else if (w.size() >= 2 && (w[1] == "extra")) { /* extra bond per atom etc. */ }
      else fail(AWSEM_E_FORMAT, "Unknown identifier in data file: " + t);
    }
    if (natoms <= 0 || ntypes <= 0) fail(AWSEM_E_FORMAT, "Data file " + file + " defines no atoms or atom types");
    mass.assign(ntypes + 1, 0.0);
    bond_k.assign(nbtypes + 1, 0.0);
    bond_r0.assign(nbtypes + 1, 0.0);
    lam.assign((size_t)(ntypes + 1) * (ntypes + 1), 0.0);
    pcs = lam;
    pcl = lam;
    pset.assign(lam.size(), 0);
    std::vector<char> mass_set(ntypes + 1, 0);
    auto rows = [&](long long count, const std::function<void(const std::vector<std::string> &)> &take) {
      long long got = 0;
      while (got < count && std::getline(in, line)) {
        std::string t = line.substr(0, line.find('#'));
        std::istringstream ls(t);
        std::vector<std::string> w;
        for (std::string s; ls >> s;) w.push_back(s);
        if (w.empty()) continue;
        take(w);
        ++got;
      }
      if (got < count) fail(AWSEM_E_FORMAT, "Unexpected end of data file in section " + section);
    };
    while (!section.empty()) {
      if (section == "Masses") {
        rows(ntypes, [&](const std::vector<std::string> &w) {
          if (w.size() < 2) fail(AWSEM_E_FORMAT, "Invalid Masses line");
          int t = (int)to_ll(w[0], "mass type");
          if (t < 1 || t > ntypes) fail(AWSEM_E_FORMAT, "Invalid type for mass set");
          mass[t] = to_d(w[1], "mass");
          mass_set[t] = 1;
        });
      } else if (section.compare(0, 5, "Atoms") == 0) {
        atoms.clear();
        rows(natoms, [&](const std::vector<std::string> &w) {
          // atom_vec_awsemmd.cpp:56: id molecule residue type q x y z [ix iy iz]
          if (w.size() != 8 && w.size() != 11) fail(AWSEM_E_FORMAT, "Incorrect atom format in data file");
          Atom A{};
          A.id = to_ll(w[0], "atom id");
          A.mol = (int)to_ll(w[1], "molecule");
          A.res = (int)to_ll(w[2], "residue");
          A.type = (int)to_ll(w[3], "atom type");
          A.q = to_d(w[4], "charge");
          for (int d = 0; d < 3; ++d) A.x[d] = to_d(w[5 + d], "coordinate");
          if (w.size() == 11)
            for (int d = 0; d < 3; ++d)
              if (to_ll(w[8 + d], "image flag") != 0) unsupported("non-zero image flags in the data file");
          if (A.type < 1 || A.type > ntypes) fail(AWSEM_E_FORMAT, "Invalid atom type in Atoms section of data file");
          if (index_of.count(A.id)) fail(AWSEM_E_FORMAT, "Duplicate atom IDs exist");
          index_of[A.id] = (int)atoms.size();
          atoms.push_back(A);
        });
      } else if (section == "Velocities") {
        rows(natoms, [&](const std::vector<std::string> &w) {
          if (w.size() < 4) fail(AWSEM_E_FORMAT, "Incorrect velocity format in data file");
          auto it = index_of.find(to_ll(w[0], "atom id"));
          if (it == index_of.end()) fail(AWSEM_E_FORMAT, "Invalid atom ID in Velocities section of data file");
          for (int d = 0; d < 3; ++d) atoms[it->second].v[d] = to_d(w[1 + d], "velocity");
        });
      } else if (section == "Bond Coeffs") {
        rows(nbtypes, [&](const std::vector<std::string> &w) {
          if (w.size() < 3) fail(AWSEM_E_FORMAT, "Invalid Bond Coeffs line");
          int t = (int)to_ll(w[0], "bond type");
          if (t < 1 || t > nbtypes) fail(AWSEM_E_FORMAT, "Invalid bond type in Bond Coeffs");
          bond_k[t] = to_d(w[1], "bond K");
          bond_r0[t] = to_d(w[2], "bond r0");
        });
      } else if (section == "Bonds") {
        rows(nbonds, [&](const std::vector<std::string> &w) {
          if (w.size() < 4) fail(AWSEM_E_FORMAT, "Incorrect format of Bonds section in data file");
          int t = (int)to_ll(w[1], "bond type");
          if (t < 1 || t > nbtypes) fail(AWSEM_E_FORMAT, "Invalid bond type in Bonds section of data file");
          bond_type.push_back(t);
          bond_a.push_back((int)to_ll(w[2], "bond atom"));
          bond_b.push_back((int)to_ll(w[3], "bond atom"));
        });
      } else if (section == "Pair Coeffs") {
        rows(ntypes, [&](const std::vector<std::string> &w) {
          std::vector<Tok> a{{"pair_coeff", false}, {w[0], false}, {w[0], false}};
          for (size_t k = 1; k < w.size(); ++k) a.push_back({w[k], false});
          pair_coeff(a);
        });
      } else {
        unsupported("data file section " + section);
      }
      // next section keyword
      section.clear();
      while (std::getline(in, line)) {
        std::string t = line.substr(0, line.find('#'));
        size_t b = t.find_first_not_of(" \t\r");
        if (b == std::string::npos) continue;
        t = t.substr(b);
        section = t.substr(0, t.find_last_not_of(" \t\r") + 1);
        break;
      }
    }
    if ((long long)atoms.size() != natoms) fail(AWSEM_E_FORMAT, "Data file " + file + " has no Atoms section");
    if ((long long)bond_type.size() != nbonds) fail(AWSEM_E_FORMAT, "Data file " + file + " has no Bonds section");
    for (int t = 1; t <= ntypes; ++t)
      if (!mass_set[t]) fail(AWSEM_E_FORMAT, "Not all per-type masses are set in " + file);
    groups["all"] = std::vector<char>(atoms.size(), 1);
    have_data = true;
  }

  void pair_coeff(const std::vector<Tok> &a)
  {
    // PairExcludedVolume::coeff, pev.cpp:200-232
    if (!pair_on) fail(AWSEM_E_STATE, "Pair_coeff command before pair_style is defined" + here());
    if (!have_data) fail(AWSEM_E_STATE, "Pair_coeff command before simulation box is defined" + here());
    if (a.size() != 4 && a.size() != 6) fail(AWSEM_E_FORMAT, "Incorrect args for pair coefficients" + here());
    int ilo, ihi, jlo, jhi;
    bounds(a[1].s, ntypes, ilo, ihi);
    bounds(a[2].s, ntypes, jlo, jhi);
    double l = to_d(a[3].s, "pair_coeff"), cs = pair_cs, cl = pair_cl;
    if (a.size() == 6) { cs = to_d(a[4].s, "pair_coeff"); cl = to_d(a[5].s, "pair_coeff"); }
    int count = 0;
    int nt = ntypes + 1;
    for (int i = ilo; i <= ihi; ++i)
      for (int j = std::max(jlo, i); j <= jhi; ++j) {
        lam[i * nt + j] = l; pcs[i * nt + j] = cs; pcl[i * nt + j] = cl; pset[i * nt + j] = 1;
        ++count;
      }
    if (count == 0) fail(AWSEM_E_FORMAT, "Incorrect args for pair coefficients" + here());
  }

  // velocity group create T seed [dist uniform|gaussian] [mom yes|no]
  // [rot no] [loop all|local|geom]: the per-atom draws, the momentum
  // removal and the exact rescale to T (dof = 3N - 3) as LAMMPS defines
  // them, with this library's own random stream
  void velocity(const std::vector<Tok> &a)
  {
    need(a, 3, "velocity");
    if (!have_data) fail(AWSEM_E_STATE, "Velocity command before simulation box is defined" + here());
    if (a[1].s != "all") unsupported("velocity on group " + a[1].s);
    if (a[2].s == "zero") {
      for (Atom &A : atoms) A.v[0] = A.v[1] = A.v[2] = 0.0;
      state_dirty = true;
      return;
    }
    if (a[2].s != "create") unsupported("velocity " + a[2].s);
    need(a, 5, "velocity");
    double T = to_d(a[3].s, "velocity temperature");
    long long seed = to_ll(a[4].s, "velocity seed");
    if (seed <= 0) fail(AWSEM_E_FORMAT, "Illegal velocity create command" + here());
    bool gauss = false, mom = true;
    for (size_t k = 5; k + 1 < a.size(); k += 2) {
      const std::string &key = a[k].s, &val = a[k + 1].s;
      if (key == "dist") gauss = val == "gaussian";
      else if (key == "mom") mom = val == "yes";
      else if (key == "rot") { if (val != "no") unsupported("velocity rot yes"); }
      else if (key == "loop") { /* one process: every loop mode is the same */ }
      else unsupported("velocity keyword " + key);
    }
    for (Atom &A : atoms) {
      double f = 1.0 / std::sqrt(mass[A.type]);
      for (int d = 0; d < 3; ++d) {
        double u;
        if (gauss) {
          double u1 = urand3((uint64_t)seed, (uint64_t)A.id, 2 * d), u2 = urand3((uint64_t)seed, (uint64_t)A.id, 2 * d + 1);
          u = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
        } else {
          u = urand3((uint64_t)seed, (uint64_t)A.id, d) - 0.5;
        }
        A.v[d] = u * f;
      }
    }
    if (mom) {
      double p[3] = {0, 0, 0}, m = 0.0;
      for (const Atom &A : atoms) {
        for (int d = 0; d < 3; ++d) p[d] += mass[A.type] * A.v[d];
        m += mass[A.type];
      }
      for (Atom &A : atoms)
        for (int d = 0; d < 3; ++d) A.v[d] -= p[d] / m;
    }
    double t = temperature();
    if (t > 0.0) {
      double s = std::sqrt(T / t);
      for (Atom &A : atoms)
        for (int d = 0; d < 3; ++d) A.v[d] *= s;
    }
    state_dirty = true;
  }

  // group ID id|type|molecule list: integers and ranges a:b[:c]
  void group(const std::vector<Tok> &a)
  {
    need(a, 3, "group");
    if (!have_data) fail(AWSEM_E_STATE, "Group command before simulation box is defined" + here());
    const std::string &style = a[2].s;
    if (style != "id" && style != "type" && style != "molecule") unsupported("group style " + style);
    std::vector<char> &g = groups[a[1].s];
    if (g.empty()) g.assign(atoms.size(), 0);
    std::vector<std::pair<long long, long long>> rng;
    std::vector<long long> step_of;
    for (size_t k = 3; k < a.size(); ++k) {
      const std::string &w = a[k].s;
      size_t c1 = w.find(':');
      if (c1 == std::string::npos) {
        long long v = to_ll(w, "group value");
        rng.push_back({v, v});
        step_of.push_back(1);
      } else {
        size_t c2 = w.find(':', c1 + 1);
        long long lo = to_ll(w.substr(0, c1), "group range"), hi, st = 1;
        if (c2 == std::string::npos) {
          hi = to_ll(w.substr(c1 + 1), "group range");
        } else {
          hi = to_ll(w.substr(c1 + 1, c2 - c1 - 1), "group range");
          st = to_ll(w.substr(c2 + 1), "group range");
        }
        if (st <= 0) fail(AWSEM_E_FORMAT, "Illegal range increment value" + here());
        rng.push_back({lo, hi});
        step_of.push_back(st);
      }
    }
    for (size_t i = 0; i < atoms.size(); ++i) {
      const Atom &A = atoms[i];
      long long v = style == "id" ? A.id : style == "type" ? A.type : A.mol;
      for (size_t r = 0; r < rng.size(); ++r)
        if (v >= rng[r].first && v <= rng[r].second && (v - rng[r].first) % step_of[r] == 0) { g[i] = 1; break; }
    }
  }

  void fix(const std::vector<Tok> &a)
  {
    need(a, 4, "fix");
    const std::string &id = a[1].s, &grp = a[2].s, &style = a[3].s;
    if (!groups.count(grp)) fail(AWSEM_E_FORMAT, "Could not find fix group ID " + grp + here());
    if (style == "nve") {
      if (grp != "all") unsupported("fix nve on group " + grp);
      have_nve = true;
    } else if (style == "langevin") {
      // fix ID all langevin Tstart Tstop damp seed
      need(a, 8, "fix langevin");
      if (grp != "all") unsupported("fix langevin on group " + grp);
      if (a.size() > 8) unsupported("fix langevin keyword " + a[8].s);
      have_langevin = true;
      lg_t0 = to_d(a[4].s, "fix langevin");
      lg_t1 = to_d(a[5].s, "fix langevin");
      lg_damp = to_d(a[6].s, "fix langevin");
      lg_seed = to_ll(a[7].s, "fix langevin");
      if (lg_damp <= 0.0 || lg_seed <= 0) fail(AWSEM_E_FORMAT, "Illegal fix langevin command" + here());
    } else if (style == "backbone") {
      // fb.cpp:86-100: fix ID alpha_carbons backbone beta_atoms oxygens coeff seq
      if (a.size() != 8) fail(AWSEM_E_FORMAT, "Illegal fix backbone command" + here());
      if (!fb_id.empty()) unsupported("a second fix backbone");
      if (!groups.count(a[4].s)) fail(AWSEM_E_FORMAT, "Could not find fix backbone beta atoms group ID" + here());
      if (!groups.count(a[5].s)) fail(AWSEM_E_FORMAT, "Could not find fix backbone oxygens group ID" + here());
      fb_id = id; fb_group = grp; fb_beta = a[4].s; fb_oxy = a[5].s;
      fb_coeff = resolve(a[6].s); fb_seq = resolve(a[7].s);
      create_context();
    } else if (style == "print") {
      // fix ID group print N "text" [file F | append F] [screen yes|no] [title T]
      need(a, 6, "fix print");
      PrintFix p;
      p.id = id;
      p.every = to_ll(a[4].s, "fix print");
      p.text = a[5].s;
      p.title = "# Fix print output for fix " + id;
      std::string mode = "w";
      for (size_t k = 6; k + 1 < a.size(); k += 2) {
        if (a[k].s == "file" || a[k].s == "append") { p.file = resolve(a[k + 1].s); mode = a[k].s == "file" ? "w" : "a"; }
        else if (a[k].s == "screen") p.screen = a[k + 1].s == "yes";
        else if (a[k].s == "title") p.title = a[k + 1].s;
        else unsupported("fix print keyword " + a[k].s);
      }
      if (p.every <= 0) fail(AWSEM_E_FORMAT, "Illegal fix print command" + here());
      if (!p.file.empty()) {
        p.fp = std::fopen(p.file.c_str(), mode.c_str());
        if (!p.fp) fail(AWSEM_E_FILE, "Cannot open fix print file " + p.file + here());
        if (mode == "w") std::fprintf(p.fp, "%s\n", p.title.c_str());
      }
      prints.push_back(p);
    } else {
      unsupported("fix style " + style);
    }
  }

  void unfix(const std::string &id)
  {
    for (size_t k = 0; k < prints.size(); ++k)
      if (prints[k].id == id) {
        if (prints[k].fp) std::fclose(prints[k].fp);
        prints.erase(prints.begin() + k);
        return;
      }
    unsupported("unfix of " + id + " (only fix print can be removed)");
  }

  // ---- the force field ---------------------------------------------------

  void create_context()
  {
    if (!have_data) fail(AWSEM_E_STATE, "Fix command before simulation box is defined" + here());
    int rc = awsem_create(fb_coeff.c_str(), fb_seq.c_str(), dir.c_str(), opts.precision, &ctx);
    if (rc) {
      char buf[1024] = "";
      if (ctx) awsem_last_error(ctx, buf, sizeof buf);
      std::string m = buf;
      if (ctx) { awsem_destroy(ctx); ctx = nullptr; }
      fail(rc, "fix backbone: " + (m.empty() ? std::string("context creation failed") : m) + here());
    }
    awsem_set_list_skin(ctx, neigh_skin);
    // Construct_Computational_Arrays: group codes per atom
    const std::vector<char> &gca = groups[fb_group], &gcb = groups[fb_beta], &go = groups[fb_oxy];
    size_t na = atoms.size();
    std::vector<int> g(na, -1), mol(na), res(na);
    for (size_t i = 0; i < na; ++i) {
      int k = (gca[i] ? 1 : 0) + (gcb[i] ? 1 : 0) + (go[i] ? 1 : 0);
      if (k > 1) fail(AWSEM_E_TOPOLOGY, "Atom " + std::to_string(atoms[i].id) + " is in two of the fix backbone groups");
      if (k == 0) unsupported("an atom outside the fix backbone groups (id " + std::to_string(atoms[i].id) + ")");
      g[i] = gca[i] ? 0 : gcb[i] ? 1 : 2;
      mol[i] = atoms[i].mol;
      res[i] = atoms[i].res;
    }
    ctx_check(awsem_set_topology(ctx, (int)na, g.data(), mol.data(), res.data()));
    ctx_check(awsem_n_residues(ctx, &n_res));
    atom_of.assign(3 * (size_t)n_res, -1);
    for (size_t i = 0; i < na; ++i) atom_of[(size_t)g[i] * n_res + (atoms[i].res - 1)] = (int)i;
    // atom types and masses: the engine's atom kinds (1 CA, 3 O, 4 CB,
    // 5 H-beta of glycine; masses 27, 28, 60, 2)
    for (int i = 0; i < n_res; ++i) {
      const Atom &ca = atoms[atom_of[i]], &cb = atoms[atom_of[n_res + i]], &o = atoms[atom_of[2 * n_res + i]];
      if (ca.type != 1 || o.type != 3 || (cb.type != 4 && cb.type != 5))
        unsupported("atom types other than 1 (CA), 3 (O), 4 (CB) and 5 (H-beta) in the fix backbone groups");
      if (mass[1] != 27.0 || mass[3] != 28.0 || mass[cb.type] != (cb.type == 4 ? 60.0 : 2.0))
        unsupported("masses other than 27 (CA), 28 (O), 60 (CB) and 2 (H-beta)");
    }
    if (!special_fene && pair_on) unsupported("pair_style vexcluded without special_bonds fene");
    if (boundary[0] == 'p' || boundary[1] == 'p' || boundary[2] == 'p')
      unsupported("periodic boundaries in a deck (awsem_set_box and awsem_compute_image cover them)");
    set_pair();
    set_bonds();
    // FixBackbone opens energy.log in its constructor (fb.cpp:101-106)
    std::string el = dir + "/energy.log";
    ctx_check(awsem_energy_log_open(ctx, el.c_str()));
    elog_open = true;
  }

  void ctx_check(int rc)
  {
    if (!rc) return;
    char buf[1024] = "";
    awsem_last_error(ctx, buf, sizeof buf);
    fail(rc, std::string(buf) + here());
  }

  void set_pair()
  {
    if (!pair_on) {
      ctx_check(awsem_set_vexcluded(ctx, 0, 2, nullptr, nullptr, nullptr));
      return;
    }
    // PairExcludedVolume::init_one, pev.cpp:250-265: unset pairs mix
    // geometrically from the diagonal
    int nt = ntypes + 1;
    double L[36] = {0}, CS[36] = {0}, CL[36] = {0};
    for (int i = 1; i <= ntypes; ++i)
      for (int j = i; j <= ntypes; ++j) {
        double l, cs, cl;
        if (pset[i * nt + j]) {
          l = lam[i * nt + j]; cs = pcs[i * nt + j]; cl = pcl[i * nt + j];
        } else {
          if (!pset[i * nt + i] || !pset[j * nt + j]) fail(AWSEM_E_STATE, "All pair coeffs are not set" + here());
          l = std::sqrt(lam[i * nt + i] * lam[j * nt + j]);
          cs = std::sqrt(pcs[i * nt + i] * pcs[j * nt + j]);
          cl = std::sqrt(pcl[i * nt + i] * pcl[j * nt + j]);
        }
        if (i > 5 || j > 5) {
          if (l != 0.0) unsupported("excluded-volume coefficients for atom types above 5");
          continue;
        }
        L[i * 6 + j] = L[j * 6 + i] = l;
        CS[i * 6 + j] = CS[j * 6 + i] = cs;
        CL[i * 6 + j] = CL[j * 6 + i] = cl;
      }
    ctx_check(awsem_set_vexcluded(ctx, 1, pair_p, L, CS, CL));
  }

  // the data file's bonds must be the chain bonds the engine evaluates:
  // CA(i)-CA(i+1), CA(i)-O(i), O(i)-CA(i+1), CA(i)-CB(i) or CA(i)-H-beta(i),
  // one K for all and one r0 per kind
  void set_bonds()
  {
    if (bond_type.empty()) {
      ctx_check(awsem_set_bonds(ctx, 0, 0.0, nullptr));
      return;
    }
    if (!bond_harmonic) unsupported("bonds without bond_style harmonic");
    // kind and residue of every atom
    std::vector<int> kind(atoms.size(), -1), resi(atoms.size(), -1);
    for (int g = 0; g < 3; ++g)
      for (int i = 0; i < n_res; ++i) {
        kind[atom_of[(size_t)g * n_res + i]] = g;
        resi[atom_of[(size_t)g * n_res + i]] = i;
      }
    // kind of each residue pair bond: 0 CA-CA, 1 CA-O, 2 O-CA(+1), 3 CA-CB, 4 CA-Hbeta
    auto kind_of = [&](int ia, int ib) -> int {
      int ga = kind[ia], gb = kind[ib], ra = resi[ia], rb = resi[ib];
      if (ga > gb || (ga == gb && ra > rb)) { std::swap(ga, gb); std::swap(ra, rb); std::swap(ia, ib); }
      if (ga == 0 && gb == 0 && rb == ra + 1 && atoms[ia].mol == atoms[ib].mol) return 0;
      if (ga == 0 && gb == 2 && ra == rb) return 1;
      if (ga == 0 && gb == 2 && ra == rb + 1 && atoms[ia].mol == atoms[ib].mol) return 2;
      if (ga == 0 && gb == 1 && ra == rb) return atoms[ib].type == 5 ? 4 : 3;
      return -1;
    };
    double k0 = -1.0, r0[5] = {0, 0, 0, 0, 0};
    bool seen[5] = {false, false, false, false, false};
    std::set<std::pair<int, int>> have;
    for (size_t b = 0; b < bond_type.size(); ++b) {
      auto ia = index_of.find(bond_a[b]), ib = index_of.find(bond_b[b]);
      if (ia == index_of.end() || ib == index_of.end()) fail(AWSEM_E_FORMAT, "Invalid atom ID in Bonds section of data file");
      int k = kind_of(ia->second, ib->second);
      if (k < 0) unsupported("bonds other than the chain bonds CA-CA, CA-O, O-CA, CA-CB and CA-H-beta (bond " +
                             std::to_string(b + 1) + ")");
      int t = bond_type[b];
      if (k0 < 0) k0 = bond_k[t];
      if (bond_k[t] != k0) unsupported("harmonic bonds with different K");
      if (seen[k] && r0[k] != bond_r0[t]) unsupported("bonds of one kind with different r0");
      seen[k] = true;
      r0[k] = bond_r0[t];
      have.insert({std::min(ia->second, ib->second), std::max(ia->second, ib->second)});
    }
    // every chain bond the engine adds must be in the data file
    size_t expect = 0;
    for (int i = 0; i < n_res; ++i) {
      int ca = atom_of[i], cb = atom_of[n_res + i], o = atom_of[2 * n_res + i];
      expect += 2;  // CA-O, CA-CB
      if (!have.count({std::min(ca, o), std::max(ca, o)}) || !have.count({std::min(ca, cb), std::max(ca, cb)}))
        unsupported("a chain without its CA-O or CA-CB bond (residue " + std::to_string(i + 1) + ")");
      if (i + 1 < n_res && atoms[atom_of[i + 1]].mol == atoms[ca].mol) {
        int ca1 = atom_of[i + 1];
        expect += 2;
        if (!have.count({std::min(ca, ca1), std::max(ca, ca1)}) || !have.count({std::min(o, ca1), std::max(o, ca1)}))
          unsupported("a chain without its CA-CA or O-CA bond (residues " + std::to_string(i + 1) + ", " +
                      std::to_string(i + 2) + ")");
      }
    }
    if (have.size() != expect || have.size() != bond_type.size()) unsupported("duplicate or extra bonds");
    ctx_check(awsem_set_bonds(ctx, 1, k0, r0));
  }

  // ---- running -----------------------------------------------------------

  void to_res(std::vector<double> &x, std::vector<double> &v) const
  {
    x.assign(9 * (size_t)n_res, 0.0);
    v.assign(9 * (size_t)n_res, 0.0);
    for (size_t s = 0; s < atom_of.size(); ++s)
      for (int d = 0; d < 3; ++d) {
        x[3 * s + d] = atoms[atom_of[s]].x[d];
        v[3 * s + d] = atoms[atom_of[s]].v[d];
      }
  }

  void from_res(const std::vector<double> &x, const std::vector<double> &v)
  {
    for (size_t s = 0; s < atom_of.size(); ++s)
      for (int d = 0; d < 3; ++d) {
        atoms[atom_of[s]].x[d] = x[3 * s + d];
        atoms[atom_of[s]].v[d] = v[3 * s + d];
      }
  }

  void batch_check(int rc)
  {
    if (!rc) return;
    char buf[1024] = "";
    awsem_batch_last_error(batch, buf, sizeof buf);
    fail(rc, std::string(buf) + here());
  }

  void pull()
  {
    std::vector<double> x(9 * (size_t)n_res), v(9 * (size_t)n_res);
    batch_check(awsem_batch_get_state(batch, x.data(), v.data(), nullptr, energy));
    from_res(x, v);
    energy_valid = true;
  }

  void run(const std::vector<Tok> &a)
  {
    if (a.size() < 2) fail(AWSEM_E_FORMAT, "Illegal run command" + here());
    for (size_t k = 2; k < a.size(); ++k) unsupported("run keyword " + a[k].s);
    long long n = to_ll(a[1].s, "run");
    if (n < 0) fail(AWSEM_E_FORMAT, "Illegal run command" + here());
    if (opts.max_steps > 0 && n > opts.max_steps) n = opts.max_steps;
    if (fb_id.empty()) unsupported("a run without fix backbone");
    if (!have_nve) unsupported("a run without fix nve (other integrators)");
    if (!batch) {
      int rc = awsem_batch_create(1, &ctx, opts.precision, &batch);
      if (rc) {
        char buf[1024] = "";
        awsem_last_error(ctx, buf, sizeof buf);
        fail(rc, std::string("batch: ") + buf + here());
      }
    }
    if (state_dirty) {
      std::vector<double> x, v;
      to_res(x, v);
      batch_check(awsem_batch_set_state(batch, x.data(), v.data()));
      state_dirty = false;
    }
    // setup: forces and energies at the current coordinates (LAMMPS
    // Verlet::setup; FixBackbone writes its energy.log row from it)
    batch_check(awsem_batch_compute(batch));
    pull();
    long long s0 = step, s1 = step + n;
    int64_t tfe = 0;
    ctx_check(awsem_tert_frust_every(ctx, &tfe));
    tf_every = (long long)tfe;
    ctx_check(awsem_nmer_frust_every(ctx, &tfe));
    nf_every = (long long)tfe;
    {
      int fm = 0, fc = 0, fd = 0;
      ctx_check(awsem_frag_frust_info(ctx, &fm, &fc, &fd, &tfe));
      ff_every = (long long)tfe;
    }
    outputs(s0, true, n == 0);
    double T0 = have_langevin ? lg_t0 : 0.0, T1 = have_langevin ? lg_t1 : 0.0;
    // without fix langevin the thermostat is off: an infinite damping time
    double damp = have_langevin ? lg_damp : 1e300;
    long long cur = s0;
    while (cur < s1) {
      long long next = s1;
      auto upd = [&](long long every) {
        if (every > 0) next = std::min(next, (cur / every + 1) * every);
      };
      upd(thermo_every);
      upd(tf_every);
      upd(nf_every);
      upd(ff_every);
      for (const Dump &d : dumps) upd(d.every);
      for (const PrintFix &p : prints) upd(p.every);
      // fix langevin ramps T over the run (delta = elapsed / run length);
      // each call ramps over its own steps
      awsem_langevin lg;
      lg.dt = dt;
      lg.t_start = T0 + (double)(cur - s0) / (double)n * (T1 - T0);
      lg.t_stop = T0 + (double)(next - s0) / (double)n * (T1 - T0);
      lg.damp = damp;
      lg.seed = (uint64_t)(have_langevin ? lg_seed : 1);
      lg.step0 = cur;
      batch_check(awsem_batch_run(batch, (int)(next - cur), &lg));
      cur = next;
      step = cur;
      pull();
      outputs(cur, false, cur == s1);
    }
    step = s1;
  }

  void outputs(long long s, bool setup, bool last)
  {
    (void)setup;
    // FixBackbone's energy.log: every evaluation on a thermo step
    if (elog_open && thermo_every > 0 && s % thermo_every == 0) ctx_check(awsem_energy_log_write(ctx, s, energy));
    if ((thermo_every > 0 && s % thermo_every == 0) || setup || last) thermo_line(s, setup);
    // FixBackbone's fragment, tertiary and n-mer frustratometers, in the
    // reference's order: every evaluation on a multiple of their output
    // intervals (fb.cpp:7537-7603), the setup one included
    auto due = [&](long long every) { return every > 0 && s % every == 0; };
    if (due(ff_every) || due(tf_every) || due(nf_every)) {
      std::vector<double> x, v;
      to_res(x, v);
      if (due(ff_every)) ctx_check(awsem_frag_frust_write(ctx, x.data(), s));
      if (tf_every > 0 && s % tf_every == 0) ctx_check(awsem_tert_frust_write(ctx, x.data(), s));
      if (nf_every > 0 && s % nf_every == 0) ctx_check(awsem_nmer_frust_write(ctx, x.data(), s));
    }
    for (Dump &d : dumps)
      if (s % d.every == 0 && s != d.last) dump_frame(d, s);
    for (PrintFix &p : prints)
      if (s % p.every == 0) {
        std::string t = substitute(p.text);
        if (p.fp) { std::fprintf(p.fp, "%s\n", t.c_str()); std::fflush(p.fp); }
        if (p.screen && !opts.quiet) std::printf("%s\n", t.c_str());
      }
  }

  void thermo_line(long long s, bool header)
  {
    if (!thermo_fp) {
      std::string f = dir + "/log.awsem";
      thermo_fp = std::fopen(f.c_str(), "w");
      if (!thermo_fp) fail(AWSEM_E_FILE, "Cannot open " + f);
    }
    std::string line;
    if (header) {
      for (const std::string &k : thermo_keys) line += (line.empty() ? "" : " ") + k;
      std::fprintf(thermo_fp, "%s\n", line.c_str());
      if (!opts.quiet) std::printf("%s\n", line.c_str());
      line.clear();
    }
    char b[64];
    for (const std::string &k : thermo_keys) {
      if (k == "step" || k == "atoms") std::snprintf(b, sizeof b, "%10lld", (long long)value(k, -1, 0));
      else std::snprintf(b, sizeof b, "%14.8g", k.compare(0, 2, "v_") == 0 || k.compare(0, 2, "f_") == 0
                                                    ? eval(k) : value(k, -1, 0));
      line += (line.empty() ? "" : " ") + std::string(b);
    }
    (void)s;
    std::fprintf(thermo_fp, "%s\n", line.c_str());
    std::fflush(thermo_fp);
    if (!opts.quiet) std::printf("%s\n", line.c_str());
  }

  // dump atom: id type xs ys zs (coordinates scaled to the box; a
  // shrink-wrapped dimension spans the atoms' extent at the frame)
  void dump_frame(Dump &d, long long s)
  {
    if (!d.fp) {
      d.fp = std::fopen(d.file.c_str(), "w");
      if (!d.fp) fail(AWSEM_E_FILE, "Cannot open dump file " + d.file);
    }
    const std::vector<char> &g = d.group == "all" ? groups["all"] : groups[d.group];
    std::vector<int> idx;
    for (size_t i = 0; i < atoms.size(); ++i) if (g[i]) idx.push_back((int)i);
    if (d.sort_id) std::sort(idx.begin(), idx.end(), [&](int p, int q) { return atoms[p].id < atoms[q].id; });
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
      if (boundary[k] == 's' || boundary[k] == 'm') {
        lo[k] = 1e300; hi[k] = -1e300;
        for (const Atom &A : atoms) { lo[k] = std::min(lo[k], A.x[k]); hi[k] = std::max(hi[k], A.x[k]); }
      } else {
        lo[k] = box[k][0]; hi[k] = box[k][1];
      }
    }
    std::fprintf(d.fp, "ITEM: TIMESTEP\n%lld\nITEM: NUMBER OF ATOMS\n%zu\nITEM: BOX BOUNDS %c%c %c%c %c%c\n", s,
                 idx.size(), boundary[0], boundary[0], boundary[1], boundary[1], boundary[2], boundary[2]);
    for (int k = 0; k < 3; ++k) std::fprintf(d.fp, "%-1.16e %-1.16e\n", lo[k], hi[k]);
    std::fprintf(d.fp, "ITEM: ATOMS id type xs ys zs\n");
    for (int i : idx) {
      const Atom &A = atoms[i];
      double f[3];
      for (int k = 0; k < 3; ++k) f[k] = hi[k] > lo[k] ? (A.x[k] - lo[k]) / (hi[k] - lo[k]) : 0.0;
      std::fprintf(d.fp, "%lld %d %g %g %g\n", A.id, A.type, f[0], f[1], f[2]);
    }
    std::fflush(d.fp);
    d.last = s;
  }

  bool run_all()
  {
    for (;;) {
      if (!exec_until_run()) return true;
      std::vector<Tok> a = tokenize(lines[pc]);
      ++pc;
      run(a);
    }
  }
};

extern "C" {

int awsem_deck_open(const char *input_path, const awsem_deck_opts *opts, awsem_deck **out)
{
  if (!input_path || !out) return AWSEM_E_ARG;
  *out = nullptr;
  std::unique_ptr<awsem_deck> d(new awsem_deck());
  if (opts) d->opts = *opts;
  try {
    d->load(input_path);
    d->exec_until_run();
  } catch (const DeckErr &e) {
    d->err = e.msg;
    d->broken = true;
    *out = d.release();
    return e.code;
  }
  *out = d.release();
  return AWSEM_OK;
}

int awsem_deck_run(awsem_deck *d)
{
  if (!d) return AWSEM_E_ARG;
  if (d->broken) return AWSEM_E_STATE;
  try {
    d->run_all();
  } catch (const DeckErr &e) {
    d->err = e.msg;
    d->broken = true;
    return e.code;
  }
  return AWSEM_OK;
}

int awsem_deck_natoms(const awsem_deck *d, int *n_atoms, int *n_residues)
{
  if (!d) return AWSEM_E_ARG;
  if (n_atoms) *n_atoms = (int)d->atoms.size();
  if (n_residues) *n_residues = d->n_res;
  return AWSEM_OK;
}

int awsem_deck_state(const awsem_deck *d, int64_t *ids, double *x, double *v, double *energy, int64_t *step)
{
  if (!d) return AWSEM_E_ARG;
  // atoms in ascending id order
  std::vector<int> idx(d->atoms.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  std::sort(idx.begin(), idx.end(), [&](int p, int q) { return d->atoms[p].id < d->atoms[q].id; });
  for (size_t k = 0; k < idx.size(); ++k) {
    const Atom &A = d->atoms[idx[k]];
    if (ids) ids[k] = A.id;
    for (int c = 0; c < 3; ++c) {
      if (x) x[3 * k + c] = A.x[c];
      if (v) v[3 * k + c] = A.v[c];
    }
  }
  if (energy) for (int k = 0; k < AWSEM_N_ET; ++k) energy[k] = d->energy[k];
  if (step) *step = d->step;
  return AWSEM_OK;
}

int awsem_deck_residue_atoms(const awsem_deck *d, int64_t *ids)
{
  if (!d || !ids) return AWSEM_E_ARG;
  if (d->atom_of.empty()) return AWSEM_E_STATE;
  for (size_t s = 0; s < d->atom_of.size(); ++s) ids[s] = d->atoms[d->atom_of[s]].id;
  return AWSEM_OK;
}

int awsem_deck_variable(awsem_deck *d, const char *expr, double *value)
{
  if (!d || !expr || !value) return AWSEM_E_ARG;
  try {
    *value = d->eval(expr);
  } catch (const DeckErr &e) {
    d->err = e.msg;
    return e.code;
  }
  return AWSEM_OK;
}

int awsem_deck_last_error(const awsem_deck *d, char *buf, size_t n)
{
  if (!d || !buf || n == 0) return AWSEM_E_ARG;
  std::snprintf(buf, n, "%s", d->err.c_str());
  return AWSEM_OK;
}

int awsem_deck_destroy(awsem_deck *d)
{
  delete d;
  return AWSEM_OK;
}

}  // extern "C"
