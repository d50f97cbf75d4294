// CHEMKIN reader: mechanism (ELEMENTS / SPECIES / REACTIONS), NASA-7
// thermodynamic data and Lennard-Jones transport data, producing the model
// constants of ostc_param exactly as the reference's Species / Reaction
// constructors do (src/ChemData.cpp:11-614, atomic weights
// src/Element.cpp:13-132).  Reference behaviours kept for parity:
//  * species index = order of appearance in the SPECIES section;
//  * the global T_low/T_mid/T_high line of the THERMO section does not change
//    the per-species defaults (300/1000/5000 K) -- only each species' own
//    columns do;
//  * the main-line activation energy is converted from cal/mol to K (/R_c);
//    the LOW and REV activation energies are stored as read;
//  * a DUPLICATE pair merges into the slot of its first member (second set
//    in A2/b2/Ea2, type 3) and the second slot is reused;
//  * third-body efficiencies default to 1; "(+M)" -> type 5, TROE with 3/4
//    parameters -> 6/7, REV -> 2, "=>" -> 0 (irreversible).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ostc.h"

namespace {

constexpr double kRc = 1.98720425864083;

std::string ltrim(const std::string &s) {
  const auto p = s.find_first_not_of(" \t\r\n\v\f");
  return p == std::string::npos ? std::string() : s.substr(p);
}
std::string upper(std::string s) {
  for (auto &c : s) c = (char)std::toupper((unsigned char)c);
  return s;
}
// a line with leading blanks removed and letters upper-cased
bool next_line(std::ifstream &f, std::string &out, bool up = true) {
  if (!std::getline(f, out)) return false;
  out = ltrim(out);
  if (up) out = upper(out);
  return true;
}

double atomic_weight(const std::string &el) {
  static const std::map<std::string, double> w = {
      {"H", 1.008},       {"HE", 4.002602}, {"LI", 6.94},   {"BE", 9.0121831}, {"B", 10.81},   {"C", 12.011},
      {"N", 14.007},      {"O", 15.999},    {"F", 18.998403163}, {"NE", 20.1797}, {"NA", 22.98976928},
      {"MG", 24.305},     {"AL", 26.9815384}, {"SI", 28.085}, {"P", 30.973761998}, {"S", 32.06}, {"CL", 35.45},
      {"AR", 39.95},      {"K", 39.0983}};
  auto it = w.find(el);
  return it == w.end() ? 0.0 : it->second;
}

struct Mech {
  std::vector<std::string> elements;     // file order
  std::map<std::string, int> elem_index;  // name -> column (file order)
  std::map<std::string, int> species;    // name -> index (file order)
  std::vector<std::string> spec_names;
  std::vector<std::vector<int>> comp;  // [species][element]
  std::string err;
};

bool read_species_section(const std::string &path, Mech &m) {
  std::ifstream f(path);
  if (!f) {
    m.err = "cannot open mechanism file " + path;
    return false;
  }
  std::string tok;
  while (f >> tok) {
    if (tok[0] == '!') {
      std::getline(f, tok);
      continue;
    }
    if (tok == "ELEMENTS" || tok == "ELEM") break;
  }
  while (f >> tok) {
    if (tok[0] == '!') {
      std::getline(f, tok);
      continue;
    }
    tok = upper(tok);
    if (tok == "END") continue;
    if (tok == "SPECIES" || tok == "SPEC") break;
    if (!m.elem_index.count(tok)) {
      m.elem_index[tok] = (int)m.elements.size();
      m.elements.push_back(tok);
    }
  }
  while (f >> tok) {
    if (tok[0] == '!') {
      std::getline(f, tok);
      continue;
    }
    tok = upper(tok);
    if (tok == "END") continue;
    if (tok == "REACTIONS" || tok == "REAC" || tok == "THERMO") break;
    if (!m.species.count(tok)) {
      m.species[tok] = (int)m.spec_names.size();
      m.spec_names.push_back(tok);
    }
  }
  const int ns = (int)m.spec_names.size();
  if (ns == 0 || ns > OSTC_MAX_SPEC) {
    m.err = "species count out of range";
    return false;
  }
  m.comp.assign(ns, std::vector<int>(m.elements.size(), 0));
  return true;
}

bool read_thermo(const std::string &path, Mech &m, ostc_param *p) {
  std::ifstream f(path);
  if (!f) {
    m.err = "cannot open thermo file " + path;
    return false;
  }
  const int ns = (int)m.spec_names.size();
  for (int s = 0; s < ns; ++s) {
    p->t_low[s] = 300;
    p->t_mid[s] = 1000;
    p->t_high[s] = 5000;
  }
  std::string line;
  while (next_line(f, line))
    if (line.rfind("THERMO", 0) == 0) break;
  while (std::getline(f, line))  // the global temperature line (no effect on the defaults)
    if (!line.empty() && line[0] != '!') break;
  std::vector<bool> done(ns, false);
  int n_read = 0;
  while (next_line(f, line)) {
    if (line.empty() || line[0] == '!') continue;
    std::istringstream ls(line);
    std::string key;
    ls >> key;
    if (key == "END") continue;
    if (key == "REACTIONS" || key == "TRANSPORT" || key == "TRAN") break;
    if (n_read >= ns) continue;
    std::istringstream ns_(line.substr(0, 18));
    std::string name;
    ns_ >> name;
    auto it = m.species.find(name);
    std::string l2, l3, l4;
    if (it == m.species.end() || done[it->second]) {
      std::getline(f, l2);
      std::getline(f, l3);
      std::getline(f, l4);
      if (it != m.species.end()) next_line(f, l4);  // the reference also consumes one more line here
      continue;
    }
    const int s = it->second;
    auto field = [&](size_t pos, size_t len) { return pos < line.size() ? line.substr(pos, len) : std::string(); };
    p->t_low[s] = std::stod(field(45, 10));
    p->t_high[s] = std::stod(field(55, 10));
    {
      std::istringstream tm(field(65, 10));
      std::string v;
      tm >> v;
      if (!v.empty()) p->t_mid[s] = std::stod(v);
    }
    for (int q = 0; q < 4; ++q) {
      std::string c = ltrim(field(24 + 5 * q, 5));
      if (c.empty() || c[0] == '0') break;
      std::istringstream cs(c);
      std::string el;
      int cnt = 0;
      cs >> el >> cnt;
      auto e = m.elem_index.find(el);
      if (e == m.elem_index.end()) {  // unknown symbols map to column 0, as std::map::operator[] would
        m.elem_index[el] = 0;
        e = m.elem_index.find(el);
      }
      m.comp[s][e->second] = cnt;
    }
    // molecular weight: elements summed in name order (std::map iteration)
    double mw = 0;
    std::map<std::string, int> sorted(m.elem_index.begin(), m.elem_index.end());
    for (const auto &kv : sorted) mw += atomic_weight(kv.first) * m.comp[s][kv.second];
    p->mw[s] = mw;
    std::getline(f, l2);
    std::getline(f, l3);
    std::getline(f, l4);
    auto num = [](const std::string &l, int q) { return std::stod(l.substr(15 * q, 15)); };
    double *hi = &p->high_temp_coeff[s * 7], *lo = &p->low_temp_coeff[s * 7];
    for (int q = 0; q < 5; ++q) hi[q] = num(l2, q);
    hi[5] = num(l3, 0);
    hi[6] = num(l3, 1);
    lo[0] = num(l3, 2);
    lo[1] = num(l3, 3);
    lo[2] = num(l3, 4);
    lo[3] = num(l4, 0);
    lo[4] = num(l4, 1);
    lo[5] = num(l4, 2);
    lo[6] = num(l4, 3);
    done[s] = true;
    ++n_read;
  }
  if (n_read < ns) {
    m.err = "thermodynamic data missing for some species";
    return false;
  }
  return true;
}

bool read_transport(const std::string &path, Mech &m, ostc_param *p) {
  std::ifstream f(path);
  if (!f) {
    m.err = "cannot open transport file " + path;
    return false;
  }
  const int ns = (int)m.spec_names.size();
  std::vector<bool> done(ns, false);
  int n_read = 0;
  std::string line;
  while (next_line(f, line)) {
    if (line.empty() || line[0] == '!') continue;
    std::istringstream ls(line);
    std::string key;
    ls >> key;
    if (key.rfind("END", 0) == 0 || key.rfind("REAC", 0) == 0) break;
    auto it = m.species.find(key);
    if (it == m.species.end()) continue;
    if (n_read >= ns) break;
    if (done[it->second]) continue;
    double geom = 0, eps = 0, sigma = 0;
    ls >> geom >> eps >> sigma;
    const int s = it->second;
    p->LJ_potent_inv[s] = 1.0 / eps;
    p->vis_coeff[s] = 2.6693e-6 * std::sqrt(p->mw[s]) / (sigma * sigma);
    done[s] = true;
    ++n_read;
  }
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j) {
      p->WjDivWi_to_One4th[i * OSTC_MAX_SPEC + j] = std::pow(p->mw[j] / p->mw[i], 0.25);
      p->sqrt_WiDivWjPl1Mul8[i * OSTC_MAX_SPEC + j] = 1.0 / std::sqrt(8 * (1 + p->mw[i] / p->mw[j]));
    }
  if (n_read < ns) {
    m.err = "transport data missing for some species";
    return false;
  }
  return true;
}

// ------------------------------------------------------------- reactions

struct Reac {
  int type = 1, order = 0;
  std::vector<int> nf, nb;
  double A = 0, b = 0, Ea = 0, A2 = 0, b2 = 0, Ea2 = 0;
  std::vector<double> tb;
  double alpha = 0, t3 = 0, t1 = 0, t2 = 0;
};

// one "A+B=C+D  A b Ea" line (upper-cased, left-trimmed)
bool parse_equation(std::string line, Reac &r, const Mech &m, std::string &err) {
  const auto lp = line.find("(+");
  if (line.find("+M") != std::string::npos) {
    r.type = 4;
    if (lp != std::string::npos) {
      r.type = 5;
      const auto rp = line.find(')', lp);
      line.erase(lp, rp - lp + 1);
      const auto lp2 = line.find("(+");
      const auto rp2 = line.find_last_of(')');
      if (lp2 != std::string::npos && rp2 != std::string::npos) line.erase(lp2, rp2 - lp2 + 1);
    } else {
      line.erase(line.find("+M"), 2);
      line.erase(line.find("+M"), 2);
    }
  }
  auto lt = line.find('<');
  if (lt == std::string::npos) lt = line.find('=');
  for (size_t q = 0; q <= lt && q < line.size(); ++q)
    if (line[q] == ' ') line[q] = '<';
  std::istringstream ls(line);
  std::string eq;
  ls >> eq >> r.A >> r.b >> r.Ea;
  r.Ea /= kRc;
  auto small = eq.find('<');
  const auto eqpos = eq.find('=');
  auto big = eq.find('>');
  if (small == std::string::npos) {
    if (big != std::string::npos)
      r.type = 0;
    else
      big = eqpos;
    small = eqpos;
  }
  std::string lhs = eq.substr(0, small), rhs = eq.substr(big + 1);
  auto side = [&](std::string s, std::vector<int> &nu, int sign) -> bool {
    while (!s.empty()) {
      auto plus = s.find('+');
      if (plus != std::string::npos && plus + 1 < s.size() && s[plus + 1] == '+') ++plus;  // '+' inside a name
      int coef = 1;
      size_t start = 0;
      if (std::isdigit((unsigned char)s[0])) {
        coef = s[0] - '0';
        start = 1;
      }
      const std::string name = plus == std::string::npos ? s.substr(start) : s.substr(start, plus - start);
      auto it = m.species.find(name);
      if (it == m.species.end()) {
        err = "unknown species in reaction: " + name;
        return false;
      }
      nu[it->second] += coef;
      r.order += sign * coef;
      if (plus == std::string::npos) break;
      s = s.substr(plus + 1);
    }
    return true;
  };
  return side(lhs, r.nf, -1) && side(rhs, r.nb, +1);
}

bool read_reactions(const std::string &path, Mech &m, ostc_param *p) {
  std::ifstream f(path);
  const int ns = (int)m.spec_names.size();
  std::string line;
  bool found = false;
  while (next_line(f, line, false))
    if (line.rfind("REACTIONS", 0) == 0 || line.rfind("REAC", 0) == 0) {
      found = true;
      break;
    }
  if (!found) {
    p->n_reac = 0;
    return true;
  }
  std::vector<Reac> reacs;
  std::vector<int> dup_open;  // indices of first members of DUPLICATE pairs
  next_line(f, line);
  while (line != "END") {
    if (line.empty() || line[0] == '!') {
      if (!next_line(f, line)) break;
      continue;
    }
    Reac r;
    r.nf.assign(ns, 0);
    r.nb.assign(ns, 0);
    r.tb.assign(ns, 1.0);
    if (!parse_equation(line, r, m, m.err)) return false;
    bool is_dup = false;
    // auxiliary lines until the next equation or END
    std::string aux;
    bool more = false;
    while ((more = next_line(f, aux))) {
      std::replace(aux.begin(), aux.end(), '/', ' ');
      std::istringstream as(aux);
      std::string key;
      as >> key;
      if (aux.empty() || aux[0] == '!') continue;
      if (key == "END" || aux.find('=') != std::string::npos) break;
      if (key == "PLOG") continue;
      if (key == "DUPLICATE" || key == "DUP") {
        is_dup = true;
        r.type = 3;
        continue;
      }
      if (key == "REV") {
        as >> r.A2 >> r.b2 >> r.Ea2;
        r.type = 2;
      } else if (key == "LOW") {
        as >> r.A2 >> r.b2 >> r.Ea2;
      } else if (key == "TROE") {
        r.type = 6;
        as >> r.alpha >> r.t3 >> r.t1;
        std::string t2;
        as >> t2;
        if (!t2.empty()) {
          r.t2 = std::stod(t2);
          r.type = 7;
        }
      } else {
        std::istringstream tbs(aux);
        std::string sp;
        double eff = 0;
        while (tbs >> sp >> eff) {
          auto it = m.species.find(sp);
          if (it == m.species.end()) {
            m.err = "unknown third-body species " + sp;
            return false;
          }
          r.tb[it->second] = eff;
        }
      }
    }
    if (!more) aux = "END";
    if (is_dup) {
      int match = -1;
      for (size_t q = 0; q < dup_open.size(); ++q)
        if (reacs[dup_open[q]].nf == r.nf && reacs[dup_open[q]].nb == r.nb) {
          match = (int)q;
          break;
        }
      if (match >= 0) {
        Reac &first = reacs[dup_open[match]];
        first.A2 = r.A;
        first.b2 = r.b;
        first.Ea2 = r.Ea;
        dup_open.erase(dup_open.begin() + match);
      } else {
        dup_open.push_back((int)reacs.size());
        reacs.push_back(r);
      }
    } else {
      reacs.push_back(r);
    }
    line = aux;
  }
  const int nr = (int)reacs.size();
  if (nr > OSTC_MAX_REAC) {
    m.err = "too many reactions";
    return false;
  }
  p->n_reac = nr;
  for (int i = 0; i < nr; ++i) {
    const Reac &r = reacs[i];
    p->reac_type[i] = r.type;
    p->reac_order[i] = r.order;
    p->A[i] = r.A;
    p->b[i] = r.b;
    p->Ea[i] = r.Ea;
    p->A2[i] = r.A2;
    p->b2[i] = r.b2;
    p->Ea2[i] = r.Ea2;
    p->troe_alpha[i] = r.alpha;
    p->troe_t3[i] = r.t3;
    p->troe_t1[i] = r.t1;
    p->troe_t2[i] = r.t2;
    for (int s = 0; s < ns; ++s) {
      p->stoi_f[i * OSTC_MAX_SPEC + s] = r.nf[s];
      p->stoi_b[i * OSTC_MAX_SPEC + s] = r.nb[s];
      p->third_body_coeff[i * OSTC_MAX_SPEC + s] = r.tb[s];
    }
  }
  return true;
}

thread_local std::string g_chem_err;

}  // namespace

extern "C" int ostc_read_chemkin(const char *mechanism, const char *thermo, const char *transport, int with_reactions,
                                 ostc_param *param, ostc_mechanism_info *info) {
  if (!mechanism || !thermo || !transport || !param) return OSTC_ERR_ARG;
  Mech m;
  ostc_param &p = *param;
  bool ok = read_species_section(mechanism, m) && read_thermo(thermo, m, &p) && read_transport(transport, m, &p);
  if (ok && with_reactions) ok = read_reactions(mechanism, m, &p);
  if (!ok) {
    g_chem_err = m.err;
    return OSTC_ERR_ARG;
  }
  if (!with_reactions) p.n_reac = 0;
  p.n_spec = (int)m.spec_names.size();
  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->n_spec = p.n_spec;
    info->n_elem = (int)std::min<size_t>(m.elements.size(), OSTC_MAX_ELEM);
    for (int s = 0; s < p.n_spec; ++s) {
      std::strncpy(info->spec_names[s], m.spec_names[s].c_str(), sizeof(info->spec_names[s]) - 1);
      for (int e = 0; e < info->n_elem; ++e) info->elem_comp[s * OSTC_MAX_ELEM + e] = m.comp[s][e];
    }
    for (int e = 0; e < info->n_elem; ++e)
      std::strncpy(info->elem_names[e], m.elements[e].c_str(), sizeof(info->elem_names[e]) - 1);
  }
  return OSTC_OK;
}

extern "C" const char *ostc_chemkin_error(void) { return g_chem_err.c_str(); }
