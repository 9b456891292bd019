#include "mechanism.hpp"
#include "fastmath.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <regex>
#include <set>
#include <sstream>
#include <stdexcept>

#include "cti_lite.hpp"
#include "xml_lite.hpp"
#include "yaml_lite.hpp"

namespace hpo {

namespace {

// Atomic weights as Cantera 2.x tabulates them (kg/kmol); pyJac's molecular weights
// (e.g. 2.016 for H2, 39.95 for AR in dydt.c:44-132) follow from these.
double atomic_weight(const std::string& e) {
  static const std::map<std::string, double> aw = {
      {"H", 1.008},   {"D", 2.014102}, {"He", 4.002602},      {"C", 12.011},  {"N", 14.007},
      {"O", 15.999},  {"F", 18.998403163}, {"Ne", 20.1797},    {"Si", 28.085}, {"S", 32.06},
      {"Cl", 35.45},  {"Ar", 39.95},   {"Kr", 83.798},        {"Xe", 131.293}, {"E", 0.000548579909}};
  auto it = aw.find(e);
  if (it == aw.end()) throw std::runtime_error("unknown element '" + e + "'");
  return it->second;
}

std::string trim(const std::string& s) {
  size_t b = s.find_first_not_of(" \t");
  if (b == std::string::npos) return "";
  size_t e = s.find_last_not_of(" \t");
  return s.substr(b, e - b + 1);
}

struct Side {
  std::vector<std::pair<std::string, double>> terms;
  std::string collider;  // "(+X)" falloff collider, empty if none
  bool plus_m = false;
};

Side parse_side(std::string s) {
  Side out;
  size_t lp = s.find("(+");
  if (lp != std::string::npos) {
    size_t rp = s.find(')', lp);
    if (rp == std::string::npos) throw std::runtime_error("bad falloff collider in '" + s + "'");
    out.collider = trim(s.substr(lp + 2, rp - lp - 2));
    s = s.substr(0, lp) + s.substr(rp + 1);
  }
  size_t pos = 0;
  while (pos <= s.size()) {
    size_t nx = s.find(" + ", pos);
    std::string tok = trim(s.substr(pos, nx == std::string::npos ? std::string::npos : nx - pos));
    if (!tok.empty()) {
      double coef = 1.0;
      std::string name = tok;
      size_t sp = tok.find(' ');
      if (sp != std::string::npos) {
        std::string c = tok.substr(0, sp);
        char* end = nullptr;
        double v = std::strtod(c.c_str(), &end);
        if (end && *end == '\0') {
          coef = v;
          name = trim(tok.substr(sp + 1));
        }
      }
      if (name == "M") out.plus_m = true;
      else out.terms.emplace_back(name, coef);
    }
    if (nx == std::string::npos) break;
    pos = nx + 3;
  }
  return out;
}

// "A + B <=> C" / "A => B" / "A = B" -> the two sides; false for an irreversible arrow.
bool split_equation(const std::string& eq, std::string& lhs, std::string& rhs) {
  size_t p;
  if ((p = eq.find("<=>")) != std::string::npos) {
    lhs = eq.substr(0, p);
    rhs = eq.substr(p + 3);
    return true;
  }
  if ((p = eq.find("=>")) != std::string::npos) {
    lhs = eq.substr(0, p);
    rhs = eq.substr(p + 2);
    return false;
  }
  if ((p = eq.find('=')) != std::string::npos) {
    lhs = eq.substr(0, p);
    rhs = eq.substr(p + 1);
    return true;
  }
  throw std::runtime_error("bad equation '" + eq + "'");
}

double pressure_factor(const std::string& u) {
  if (u == "Pa") return 1.0;
  if (u == "kPa") return 1e3;
  if (u == "MPa") return 1e6;
  if (u == "bar") return 1e5;
  if (u == "atm") return kPAtm;
  if (u == "torr") return kPAtm / 760.0;
  throw std::runtime_error("pressure unit " + u);
}

// Sorts the levels by pressure; a pressure given twice (a sum of rates at one level) and a
// non-positive pre-exponential factor are refused.
void finish_plog(Reaction& rx) {
  std::sort(rx.plog.begin(), rx.plog.end(), [](const PlogLevel& a, const PlogLevel& b) { return a.lnP < b.lnP; });
  for (size_t i = 0; i < rx.plog.size(); ++i) {
    if (!std::isfinite(rx.plog[i].lnA)) throw std::runtime_error("PLOG with a non-positive A unsupported: " + rx.equation);
    if (i > 0 && rx.plog[i].lnP == rx.plog[i - 1].lnP)
      throw std::runtime_error("PLOG with several rates at one pressure unsupported: " + rx.equation);
  }
  if (rx.plog.empty()) throw std::runtime_error("PLOG without pressure levels: " + rx.equation);
  rx.lnA = rx.plog[0].lnA;
  rx.A = rx.plog[0].A;
  rx.b = rx.plog[0].b;
  rx.theta = rx.plog[0].theta;
}

std::vector<std::pair<int, double>> merge(const std::vector<std::pair<std::string, double>>& t,
                                          const std::map<std::string, int>& index) {
  std::map<int, double> acc;
  for (auto& kv : t) {
    auto it = index.find(kv.first);
    if (it == index.end()) throw std::runtime_error("unknown species '" + kv.first + "' in reaction");
    acc[it->second] += kv.second;
  }
  return {acc.begin(), acc.end()};
}

double unit_factor_conc(const YNode& units) {
  // (length^3 / quantity) -> m^3/kmol
  double lf = 1.0, qf = 1.0;
  if (const YNode* l = units.find("length")) {
    const std::string& u = l->as_string();
    lf = u == "cm" ? 1e-2 : u == "mm" ? 1e-3 : u == "m" ? 1.0 : throw std::runtime_error("length unit " + u);
  }
  if (const YNode* q = units.find("quantity")) {
    const std::string& u = q->as_string();
    qf = u == "mol" ? 1e-3 : u == "kmol" ? 1.0 : u == "molec" ? 1.0 / 6.02214076e26
                                                : throw std::runtime_error("quantity unit " + u);
  }
  return lf * lf * lf / qf;
}

double unit_factor_ea(const YNode& units) {
  // J/kmol per activation-energy unit
  const YNode* a = units.find("activation-energy");
  if (!a) return 1.0;
  const std::string& u = a->as_string();
  if (u == "cal/mol") return 4184.0;
  if (u == "kcal/mol") return 4184.0e3;
  if (u == "J/mol") return 1e3;
  if (u == "kJ/mol") return 1e6;
  if (u == "J/kmol") return 1.0;
  if (u == "K") return kRGas;
  throw std::runtime_error("activation-energy unit " + u);
}

void arrhenius(const YNode& rc, double order, double cf, double ef, double& lnA, double& b, double& th,
               double& A) {
  A = rc["A"].as_double() * std::pow(cf, order - 1.0);
  b = rc["b"].as_double();
  th = rc["Ea"].as_double() * ef / kRGas;
  lnA = A > 0 ? std::log(A) : -INFINITY;
}

}  // namespace

Mechanism load_mechanism_yaml(const std::string& path) {
  YNode doc = yaml_load_file(path);
  Mechanism m;
  m.path = path;
  YNode nounits;
  nounits.kind = YNode::Map;
  const YNode& units = doc.has("units") ? doc["units"] : nounits;
  const YNode& phase = doc["phases"][0];
  const YNode& spl = phase["species"];
  std::map<std::string, int> index;
  for (size_t i = 0; i < spl.size(); ++i) {
    m.species.push_back(spl[i].as_string());
    index[spl[i].as_string()] = static_cast<int>(i);
  }
  const int K = m.K();
  m.W.assign(K, 0.0);
  m.tmid.assign(K, 0.0);
  m.nasa.assign(static_cast<size_t>(K) * 14, 0.0);
  const YNode& sdb = doc["species"];
  std::map<std::string, const YNode*> byname;
  for (size_t i = 0; i < sdb.size(); ++i) byname[sdb[i]["name"].as_string()] = &sdb[i];
  for (int k = 0; k < K; ++k) {
    auto it = byname.find(m.species[k]);
    if (it == byname.end()) throw std::runtime_error("species '" + m.species[k] + "' has no entry");
    const YNode& s = *it->second;
    double w = 0;
    for (auto& kv : s["composition"].map) w += atomic_weight(kv.first) * kv.second.as_double();
    m.W[k] = w;
    const YNode& th = s["thermo"];
    if (th["model"].as_string() != "NASA7")
      throw std::runtime_error(m.species[k] + ": thermo model " + th["model"].as_string() + " unsupported");
    const YNode& tr = th["temperature-ranges"];
    const YNode& data = th["data"];
    const YNode& lo = data[0];
    const YNode& hi = data.size() > 1 ? data[1] : data[0];
    m.tmid[k] = data.size() > 1 ? tr[1].as_double() : tr[tr.size() - 1].as_double();
    for (int j = 0; j < 7; ++j) {
      m.nasa[14 * k + j] = lo[j].as_double();
      m.nasa[14 * k + 7 + j] = hi[j].as_double();
    }
  }
  const double cf = unit_factor_conc(units), ef = unit_factor_ea(units);
  const YNode& rl = doc["reactions"];
  for (size_t ri = 0; ri < rl.size(); ++ri) {
    const YNode& rd = rl[ri];
    Reaction rx;
    rx.equation = rd["equation"].as_string();
    std::string eq = rx.equation, lhs, rhs;
    rx.reversible = split_equation(eq, lhs, rhs);
    if (rd.has("orders")) throw std::runtime_error("non-standard reaction orders unsupported: " + eq);
    Side L = parse_side(lhs), Rr = parse_side(rhs);
    rx.reactants = merge(L.terms, index);
    rx.products = merge(Rr.terms, index);
    double order = 0;
    for (auto& kv : rx.reactants) order += kv.second;
    std::string type = rd.has("type") ? rd["type"].as_string() : std::string("elementary");
    std::vector<std::pair<int, double>> effs;
    if (const YNode* e = rd.find("efficiencies"))
      for (auto& kv : e->map) {
        auto it = index.find(kv.first);
        if (it != index.end()) effs.emplace_back(it->second, kv.second.as_double());
      }
    std::sort(effs.begin(), effs.end());
    double deff = rd.has("default-efficiency") ? rd["default-efficiency"].as_double() : 1.0;
    auto set_eff = [&](double base, const std::vector<std::pair<int, double>>& e) {
      rx.eff_base = base;
      rx.eff.clear();
      for (auto& kv : e) rx.eff.emplace_back(kv.first, kv.second - base);
    };
    if (type == "three-body" || (type == "elementary" && L.plus_m)) {
      rx.kind = kThreeBody;
      arrhenius(rd["rate-constant"], order + 1, cf, ef, rx.lnA, rx.b, rx.theta, rx.A);
      set_eff(deff, effs);
    } else if (type == "falloff") {
      arrhenius(rd["high-P-rate-constant"], order, cf, ef, rx.lnA, rx.b, rx.theta, rx.A);
      arrhenius(rd["low-P-rate-constant"], order + 1, cf, ef, rx.lnA0, rx.b0, rx.theta0, rx.A0);
      if (!L.collider.empty() && L.collider != "M") {
        auto it = index.find(L.collider);
        if (it == index.end()) throw std::runtime_error("unknown collider " + L.collider);
        set_eff(0.0, {{it->second, 1.0}});
      } else {
        set_eff(deff, effs);
      }
      if (const YNode* t = rd.find("Troe")) {
        rx.kind = kTroe;
        rx.troe[0] = (*t)["A"].as_double();
        rx.troe[1] = (*t)["T3"].as_double();
        rx.troe[2] = (*t)["T1"].as_double();
        rx.troe_has_t2 = t->has("T2");
        rx.troe[3] = rx.troe_has_t2 ? (*t)["T2"].as_double() : 0.0;
      } else if (rd.has("SRI")) {
        throw std::runtime_error("SRI falloff unsupported: " + eq);
      } else {
        rx.kind = kLindemann;
      }
    } else if (type == "elementary") {
      rx.kind = kElementary;
      arrhenius(rd["rate-constant"], order, cf, ef, rx.lnA, rx.b, rx.theta, rx.A);
    } else if (type == "pressure-dependent-Arrhenius") {
      rx.kind = kElementary;
      const YNode& lv = rd["rate-constants"];
      for (size_t i = 0; i < lv.size(); ++i) {
        PlogLevel L;
        const YNode& pn = lv[i]["P"];
        double pv = 0.0;
        std::istringstream ps(pn.as_string());
        std::string unit;
        ps >> pv;
        if (ps >> unit) pv *= pressure_factor(unit);
        arrhenius(lv[i], order, cf, ef, L.lnA, L.b, L.theta, L.A);
        L.lnP = std::log(pv);
        rx.plog.push_back(L);
      }
      finish_plog(rx);
    } else {
      throw std::runtime_error("reaction type '" + type + "' unsupported: " + eq);
    }
    m.reactions.push_back(std::move(rx));
  }
  return m;
}

// ---------------------------------------------------------------------------
// Cantera CTML (XML) files, e.g. resources/mechanisms/h2_li_2004.xml of the reference.
// Rate constants are in SI-kmol units already (m, kmol, s); activation energies carry a units
// attribute. NASA7 thermo, elementary, three-body, Lindemann and Troe falloff reactions.
// ---------------------------------------------------------------------------
namespace {

std::vector<std::string> words(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string w;
  while (is >> w) out.push_back(w);
  return out;
}

std::vector<double> numbers(const std::string& s) {
  std::string t = s;
  for (char& c : t)
    if (c == ',') c = ' ';
  std::vector<double> out;
  for (const std::string& w : words(t)) out.push_back(std::stod(w));
  return out;
}

// "H:2 O:1" -> pairs
std::vector<std::pair<std::string, double>> name_values(const std::string& s) {
  std::vector<std::pair<std::string, double>> out;
  for (const std::string& w : words(s)) {
    size_t c = w.rfind(':');
    if (c == std::string::npos) throw std::runtime_error("expected name:value, got '" + w + "'");
    out.emplace_back(w.substr(0, c), std::stod(w.substr(c + 1)));
  }
  return out;
}

double ea_factor(const std::string& u) {
  if (u.empty() || u == "cal/mol") return 4184.0;
  if (u == "kcal/mol") return 4184.0e3;
  if (u == "J/mol") return 1e3;
  if (u == "kJ/mol") return 1e6;
  if (u == "J/kmol") return 1.0;
  if (u == "K") return kRGas;
  throw std::runtime_error("activation-energy unit " + u);
}

void xml_arrhenius(const XNode& a, double& lnA, double& b, double& th, double& A) {
  const XNode* xa = a.child("A");
  const XNode* xb = a.child("b");
  const XNode* xe = a.child("E");
  if (!xa) throw std::runtime_error("Arrhenius without A");
  A = std::stod(xa->text);
  b = xb ? std::stod(xb->text) : 0.0;
  th = xe ? std::stod(xe->text) * ea_factor(xe->get("units")) / kRGas : 0.0;
  lnA = A > 0 ? std::log(A) : -INFINITY;
}

const XNode* find_by_id(const XNode& root, const std::string& tag, const std::string& id) {
  for (auto& c : root.children) {
    if (c->name == tag && (id.empty() || c->get("id") == id)) return c.get();
    if (const XNode* r = find_by_id(*c, tag, id)) return r;
  }
  return nullptr;
}

}  // namespace

Mechanism load_mechanism_xml(const std::string& path) {
  std::unique_ptr<XNode> doc = xml_load_file(path);
  Mechanism m;
  m.path = path;
  const XNode* phase = find_by_id(*doc, "phase", "");
  if (!phase) throw std::runtime_error("no <phase> in " + path);
  const XNode* sa = phase->child("speciesArray");
  if (!sa) throw std::runtime_error("phase without speciesArray");
  std::map<std::string, int> index;
  for (const std::string& n : words(sa->text)) {
    index[n] = static_cast<int>(m.species.size());
    m.species.push_back(n);
  }
  std::string sd = sa->get("datasrc");
  const XNode* sdata = find_by_id(*doc, "speciesData", sd.size() > 1 && sd[0] == '#' ? sd.substr(1) : "");
  if (!sdata) throw std::runtime_error("species data '" + sd + "' not found");
  const int K = m.K();
  m.W.assign(K, 0.0);
  m.tmid.assign(K, 0.0);
  m.nasa.assign(static_cast<size_t>(K) * 14, 0.0);
  std::vector<bool> seen(K, false);
  for (const XNode* sp : sdata->all("species")) {
    auto it = index.find(sp->get("name"));
    if (it == index.end()) continue;
    const int k = it->second;
    seen[k] = true;
    double w = 0;
    if (const XNode* aa = sp->child("atomArray"))
      for (auto& kv : name_values(aa->text)) w += atomic_weight(kv.first) * kv.second;
    m.W[k] = w;
    const XNode* th = sp->child("thermo");
    if (!th) throw std::runtime_error(m.species[k] + ": no thermo");
    std::vector<const XNode*> blocks = th->all("NASA");
    if (blocks.empty()) throw std::runtime_error(m.species[k] + ": only NASA7 thermo is supported");
    std::sort(blocks.begin(), blocks.end(),
              [](const XNode* a, const XNode* b) { return std::stod(a->get("Tmin")) < std::stod(b->get("Tmin")); });
    const XNode* lo = blocks.front();
    const XNode* hi = blocks.back();
    auto coeffs = [&](const XNode* b) {
      const XNode* fa = b->child("floatArray");
      if (!fa) throw std::runtime_error(m.species[k] + ": NASA block without coefficients");
      std::vector<double> c = numbers(fa->text);
      if (c.size() != 7) throw std::runtime_error(m.species[k] + ": NASA block needs 7 coefficients");
      return c;
    };
    std::vector<double> cl = coeffs(lo), ch = coeffs(hi);
    m.tmid[k] = std::stod(lo->get("Tmax"));
    for (int j = 0; j < 7; ++j) {
      m.nasa[14 * k + j] = cl[j];
      m.nasa[14 * k + 7 + j] = ch[j];
    }
  }
  for (int k = 0; k < K; ++k)
    if (!seen[k]) throw std::runtime_error("species '" + m.species[k] + "' has no entry");

  const XNode* ra = phase->child("reactionArray");
  if (!ra) return m;
  std::string rd = ra->get("datasrc");
  const XNode* rdata = find_by_id(*doc, "reactionData", rd.size() > 1 && rd[0] == '#' ? rd.substr(1) : "");
  if (!rdata) throw std::runtime_error("reaction data '" + rd + "' not found");
  for (const XNode* x : rdata->all("reaction")) {
    Reaction rx;
    const XNode* eqn = x->child("equation");
    rx.equation = eqn ? eqn->text : x->get("id");
    rx.reversible = x->get("reversible", "yes") == "yes";
    if (x->child("order")) throw std::runtime_error("non-standard reaction orders unsupported: " + rx.equation);
    std::vector<std::pair<std::string, double>> rt, pt;
    if (const XNode* r = x->child("reactants")) rt = name_values(r->text);
    if (const XNode* p = x->child("products")) pt = name_values(p->text);
    rx.reactants = merge(rt, index);
    rx.products = merge(pt, index);
    const XNode* rc = x->child("rateCoeff");
    if (!rc) throw std::runtime_error("reaction without rateCoeff: " + rx.equation);
    const std::string type = x->get("type", "");
    const XNode* hiA = nullptr;
    const XNode* k0 = nullptr;
    for (const XNode* a : rc->all("Arrhenius")) {
      if (a->get("name") == "k0") k0 = a;
      else hiA = a;
    }
    if (!hiA) throw std::runtime_error("reaction without Arrhenius rate: " + rx.equation);
    xml_arrhenius(*hiA, rx.lnA, rx.b, rx.theta, rx.A);
    std::vector<std::pair<int, double>> effs;
    double deff = 1.0;
    if (const XNode* e = rc->child("efficiencies")) {
      deff = std::stod(e->get("default", "1.0"));
      for (auto& kv : name_values(e->text)) {
        auto it = index.find(kv.first);
        if (it != index.end()) effs.emplace_back(it->second, kv.second);
      }
    }
    std::sort(effs.begin(), effs.end());
    rx.eff_base = deff;
    for (auto& kv : effs) rx.eff.emplace_back(kv.first, kv.second - deff);
    if (type.empty() || type == "elementary") {
      rx.kind = kElementary;
    } else if (type == "plog") {
      rx.kind = kElementary;
      for (const XNode* a : rc->all("Arrhenius")) {
        const XNode* pe = a->child("P");
        if (!pe) throw std::runtime_error("PLOG rate without a pressure: " + rx.equation);
        PlogLevel L;
        xml_arrhenius(*a, L.lnA, L.b, L.theta, L.A);
        L.lnP = std::log(std::stod(pe->text) * pressure_factor(pe->get("units", "Pa")));
        rx.plog.push_back(L);
      }
      finish_plog(rx);
    } else if (type == "threeBody") {
      rx.kind = kThreeBody;
    } else if (type == "falloff") {
      if (!k0) throw std::runtime_error("falloff reaction without k0: " + rx.equation);
      xml_arrhenius(*k0, rx.lnA0, rx.b0, rx.theta0, rx.A0);
      const XNode* f = rc->child("falloff");
      const std::string ft = f ? f->get("type", "Lindemann") : "Lindemann";
      if (ft == "Troe") {
        std::vector<double> t = numbers(f->text);
        if (t.size() < 3 || t.size() > 4) throw std::runtime_error("Troe falloff needs 3 or 4 parameters");
        rx.kind = kTroe;
        rx.troe[0] = t[0];
        rx.troe[1] = t[1];
        rx.troe[2] = t[2];
        rx.troe_has_t2 = t.size() == 4;
        rx.troe[3] = rx.troe_has_t2 ? t[3] : 0.0;
      } else if (ft == "Lindemann") {
        rx.kind = kLindemann;
      } else {
        throw std::runtime_error("falloff type '" + ft + "' unsupported: " + rx.equation);
      }
    } else {
      throw std::runtime_error("reaction type '" + type + "' unsupported: " + rx.equation);
    }
    m.reactions.push_back(std::move(rx));
  }
  return m;
}

// ---------------------------------------------------------------------------
// Cantera CTI files (resources/mechanisms/*.cti of the reference): units(...), the first
// ideal_gas(...) phase, species(...) with NASA7 thermo, reaction(...), three_body_reaction(...)
// and falloff_reaction(...) with Lindemann or Troe falloff. Units follow the file's units()
// statement; without one, Cantera's SI defaults (m, kmol, s, J/kmol).
// ---------------------------------------------------------------------------
namespace {

struct CtiUnits {
  double conc = 1.0;   // (length^3 / quantity) -> m^3/kmol
  double ea = 1.0;     // activation-energy unit -> J/kmol
};

CtiUnits cti_units(const CNode* u) {
  CtiUnits out;
  if (!u) return out;
  double lf = 1.0, qf = 1.0;
  if (const CNode* l = u->kw("length")) {
    const std::string& s = l->as_string();
    lf = s == "cm" ? 1e-2 : s == "mm" ? 1e-3 : s == "m" ? 1.0 : throw std::runtime_error("length unit " + s);
  }
  if (const CNode* q = u->kw("quantity")) {
    const std::string& s = q->as_string();
    qf = s == "mol" ? 1e-3 : s == "kmol" ? 1.0 : s == "molec" ? 1.0 / 6.02214076e26
                                                  : throw std::runtime_error("quantity unit " + s);
  }
  out.conc = lf * lf * lf / qf;
  if (const CNode* e = u->kw("act_energy")) out.ea = ea_factor(e->as_string());
  if (const CNode* t = u->kw("time"))
    if (t->as_string() != "s") throw std::runtime_error("time unit " + t->as_string());
  return out;
}

void cti_arrhenius3(const CNode* a, const CNode* bb, const CNode* e, double order, const CtiUnits& u, double& lnA,
                    double& b, double& th, double& A);

// [A, b, E] (E a number or (value, unit)) or Arrhenius(A=, b=, E=), converted to SI-kmol.
void cti_arrhenius(const CNode* r, double order, const CtiUnits& u, double& lnA, double& b, double& th, double& A) {
  if (!r) throw std::runtime_error("missing rate expression");
  const CNode *a = nullptr, *bb = nullptr, *e = nullptr;
  if (r->kind == CNode::List || r->kind == CNode::Tuple) {
    a = r->arg(0);
    bb = r->arg(1);
    e = r->arg(2);
  } else if (r->kind == CNode::Call && r->str == "Arrhenius") {
    a = r->kw("A") ? r->kw("A") : r->arg(0);
    bb = r->kw("b") ? r->kw("b") : r->arg(1);
    e = r->kw("E") ? r->kw("E") : r->arg(2);
  } else {
    throw std::runtime_error("unsupported rate expression");
  }
  cti_arrhenius3(a, bb, e, order, u, lnA, b, th, A);
}

// A, b and E (a number or (value, unit)) converted to SI-kmol.
void cti_arrhenius3(const CNode* a, const CNode* bb, const CNode* e, double order, const CtiUnits& u, double& lnA,
                    double& b, double& th, double& A) {
  if (!a) throw std::runtime_error("Arrhenius without A");
  A = a->as_double() * std::pow(u.conc, order - 1.0);
  b = bb ? bb->as_double() : 0.0;
  double ef = u.ea, ev = 0.0;
  if (e && (e->kind == CNode::Tuple || e->kind == CNode::List)) {
    ev = e->arg(0)->as_double();
    if (e->arg(1)) ef = ea_factor(e->arg(1)->as_string());
  } else if (e) {
    ev = e->as_double();
  }
  th = ev * ef / kRGas;
  lnA = A > 0 ? std::log(A) : -INFINITY;
}

}  // namespace

Mechanism load_mechanism_cti(const std::string& path) {
  std::vector<std::unique_ptr<CNode>> st = cti_load_file(path);
  Mechanism m;
  m.path = path;
  const CNode* units = nullptr;
  const CNode* phase = nullptr;
  std::map<std::string, const CNode*> spdata;
  for (auto& s : st) {
    if (s->str == "units" && !units) units = s.get();
    else if (s->str == "ideal_gas" && !phase) phase = s.get();
    else if (s->str == "species")
      if (const CNode* n = s->kw("name")) spdata[n->as_string()] = s.get();
  }
  if (!phase) throw std::runtime_error("no ideal_gas phase in " + path);
  const CtiUnits u = cti_units(units);
  const CNode* spl = phase->kw("species");
  if (!spl) throw std::runtime_error("phase without species");
  std::map<std::string, int> index;
  for (const std::string& n : words(spl->as_string())) {
    if (n.find(':') != std::string::npos) throw std::runtime_error("species from other files unsupported: " + n);
    index[n] = static_cast<int>(m.species.size());
    m.species.push_back(n);
  }
  const int K = m.K();
  m.W.assign(K, 0.0);
  m.tmid.assign(K, 0.0);
  m.nasa.assign(static_cast<size_t>(K) * 14, 0.0);
  for (int k = 0; k < K; ++k) {
    auto it = spdata.find(m.species[k]);
    if (it == spdata.end()) throw std::runtime_error("species '" + m.species[k] + "' has no entry");
    const CNode& s = *it->second;
    const CNode* at = s.kw("atoms");
    double w = 0;
    if (at)
      for (auto& kv : name_values(at->as_string())) w += atomic_weight(kv.first) * kv.second;
    m.W[k] = w;
    const CNode* th = s.kw("thermo");
    if (!th) throw std::runtime_error(m.species[k] + ": no thermo");
    std::vector<const CNode*> blocks;
    if (th->kind == CNode::Call) blocks.push_back(th);
    else
      for (auto& b : th->items) blocks.push_back(b.get());
    for (const CNode* b : blocks)
      if (b->kind != CNode::Call || b->str != "NASA" || !b->arg(0) || !b->arg(1) || b->arg(0)->items.size() != 2 ||
          b->arg(1)->items.size() != 7)
        throw std::runtime_error(m.species[k] + ": only NASA7 thermo is supported");
    std::sort(blocks.begin(), blocks.end(), [](const CNode* a, const CNode* b) {
      return a->arg(0)->arg(0)->as_double() < b->arg(0)->arg(0)->as_double();
    });
    const CNode* lo = blocks.front();
    const CNode* hi = blocks.back();
    m.tmid[k] = lo->arg(0)->arg(1)->as_double();
    for (int j = 0; j < 7; ++j) {
      m.nasa[14 * k + j] = lo->arg(1)->arg(j)->as_double();
      m.nasa[14 * k + 7 + j] = hi->arg(1)->arg(j)->as_double();
    }
  }
  const CNode* rsel = phase->kw("reactions");
  if (!rsel) return m;
  if (rsel->kind != CNode::Str || rsel->str != "all")
    throw std::runtime_error("only reactions = 'all' is supported");
  for (auto& sp : st) {
    const CNode& s = *sp;
    const bool elem = s.str == "reaction", three = s.str == "three_body_reaction", fall = s.str == "falloff_reaction";
    const bool plog = s.str == "pdep_arrhenius";
    if (!elem && !three && !fall && !plog) {
      if (s.str == "chebyshev_reaction" || s.str == "chemically_activated_reaction" || s.str == "surface_reaction" ||
          s.str == "edge_reaction")
        throw std::runtime_error(s.str + " unsupported");
      continue;
    }
    Reaction rx;
    const CNode* eqn = s.kw("equation") ? s.kw("equation") : s.arg(0);
    if (!eqn) throw std::runtime_error("reaction without equation");
    rx.equation = eqn->as_string();
    if (s.kw("order")) throw std::runtime_error("non-standard reaction orders unsupported: " + rx.equation);
    std::string lhs, rhs;
    rx.reversible = split_equation(rx.equation, lhs, rhs);
    Side L = parse_side(lhs), Rr = parse_side(rhs);
    rx.reactants = merge(L.terms, index);
    rx.products = merge(Rr.terms, index);
    double order = 0;
    for (auto& kv : rx.reactants) order += kv.second;
    std::vector<std::pair<int, double>> effs;
    if (const CNode* e = s.kw("efficiencies"))
      for (auto& kv : name_values(e->as_string())) {
        auto it = index.find(kv.first);
        if (it != index.end()) effs.emplace_back(it->second, kv.second);
      }
    std::sort(effs.begin(), effs.end());
    const double deff = s.kw("default_efficiency") ? s.kw("default_efficiency")->as_double() : 1.0;
    auto set_eff = [&](double base, const std::vector<std::pair<int, double>>& e) {
      rx.eff_base = base;
      rx.eff.clear();
      for (auto& kv : e) rx.eff.emplace_back(kv.first, kv.second - base);
    };
    const CNode* kf = s.kw("kf") ? s.kw("kf") : s.arg(1);
    if (plog) {
      // each positional argument after the equation: [(P, unit), A, b, E]
      rx.kind = kElementary;
      for (size_t i = 1; i < s.items.size(); ++i) {
        const CNode* lv = s.arg(i);
        if (!lv || (lv->kind != CNode::List && lv->kind != CNode::Tuple) || lv->items.size() != 4)
          throw std::runtime_error("pdep_arrhenius level must be [(P, unit), A, b, E]: " + rx.equation);
        const CNode* pn = lv->arg(0);
        double pv = 0.0;
        if (pn->kind == CNode::Tuple || pn->kind == CNode::List)
          pv = pn->arg(0)->as_double() * (pn->arg(1) ? pressure_factor(pn->arg(1)->as_string()) : 1.0);
        else
          pv = pn->as_double();
        PlogLevel L;
        cti_arrhenius3(lv->arg(1), lv->arg(2), lv->arg(3), order, u, L.lnA, L.b, L.theta, L.A);
        L.lnP = std::log(pv);
        rx.plog.push_back(L);
      }
      finish_plog(rx);
    } else if (elem) {
      rx.kind = kElementary;
      cti_arrhenius(kf, order, u, rx.lnA, rx.b, rx.theta, rx.A);
    } else if (three) {
      rx.kind = kThreeBody;
      cti_arrhenius(kf, order + 1, u, rx.lnA, rx.b, rx.theta, rx.A);
      set_eff(deff, effs);
    } else {
      cti_arrhenius(kf, order, u, rx.lnA, rx.b, rx.theta, rx.A);
      cti_arrhenius(s.kw("kf0") ? s.kw("kf0") : s.arg(2), order + 1, u, rx.lnA0, rx.b0, rx.theta0, rx.A0);
      if (!L.collider.empty() && L.collider != "M") {
        auto it = index.find(L.collider);
        if (it == index.end()) throw std::runtime_error("unknown collider " + L.collider);
        set_eff(0.0, {{it->second, 1.0}});
      } else {
        set_eff(deff, effs);
      }
      const CNode* f = s.kw("falloff");
      if (!f || (f->kind == CNode::Call && f->str == "Lindemann")) {
        rx.kind = kLindemann;
      } else if (f->kind == CNode::Call && f->str == "Troe") {
        rx.kind = kTroe;
        const char* names[4] = {"A", "T3", "T1", "T2"};
        for (int j = 0; j < 4; ++j) {
          const CNode* v = f->kw(names[j]) ? f->kw(names[j]) : f->arg(static_cast<size_t>(j));
          if (!v && j < 3) throw std::runtime_error("Troe falloff needs A, T3 and T1: " + rx.equation);
          if (j == 3) rx.troe_has_t2 = v != nullptr;
          rx.troe[j] = v ? v->as_double() : 0.0;
        }
      } else {
        throw std::runtime_error("falloff type '" + (f->kind == CNode::Call ? f->str : std::string("?")) +
                                 "' unsupported: " + rx.equation);
      }
    }
    m.reactions.push_back(std::move(rx));
  }
  return m;
}

Mechanism load_mechanism(const std::string& path) {
  auto ends = [&](const char* e) {
    const size_t n = std::char_traits<char>::length(e);
    return path.size() >= n && path.compare(path.size() - n, n, e) == 0;
  };
  if (ends(".xml") || ends(".ctml")) return load_mechanism_xml(path);
  if (ends(".cti")) return load_mechanism_cti(path);
  return load_mechanism_yaml(path);
}

double enthalpy_mass(const Mechanism& m, double T, const double* Y) {
  double h = 0.0;
  for (int k = 0; k < m.K(); ++k) {
    const double* a = &m.nasa[14 * k + (T <= m.tmid[k] ? 0 : 7)];
    const double hRT = a[0] + T * (a[1] / 2 + T * (a[2] / 3 + T * (a[3] / 4 + T * a[4] / 5))) + a[5] / T;
    h += Y[k] * hRT / m.W[k];
  }
  return h * kGasConstant * T;
}

// ---------------------------------------------------------------------------
// Device tables
// ---------------------------------------------------------------------------
namespace {

// Troe coefficients for the device: 1/T***, 1/T* and T**, magnitudes above 1e8 held at 1e8.
// The terms they enter, exp(-T/T***), exp(-T/T*) and exp(-T**/T), are then 0 in double
// precision for every T between 1e-5 K and 1e5 K, as with the true values (nDodecane_Reitz has
// T*** = 1e-30 and T** = 1e100), and the device exponential's arguments stay within 1e13 in
// magnitude, where its argument reduction holds (an argument of -1e97 would not).
void troe_device_coefs(const Reaction& rx, double& iT3, double& iT1, double& T2) {
  constexpr double kMax = 1e8;
  iT3 = std::fmin(1.0 / rx.troe[1], kMax);
  iT1 = std::fmin(1.0 / rx.troe[2], kMax);
  T2 = std::fmin(rx.troe[3], kMax);
}

int align_to(int x, int a) { return (x + a - 1) / a * a; }

int coef_index(std::vector<double>& tab, double v) {
  for (size_t i = 0; i < tab.size(); ++i)
    if (tab[i] == v) return static_cast<int>(i);
  if (tab.size() >= 16) throw std::runtime_error("more than 15 distinct stoichiometric coefficients");
  tab.push_back(v);
  return static_cast<int>(tab.size() - 1);
}

}  // namespace

DevMechBlob build_device_tables(const Mechanism& m) {
  const int K = m.K(), R = m.R();
  if (K < 2 || K > 63) throw std::runtime_error("device path supports 2..63 species");
  if (R > 4095) throw std::runtime_error("device path supports at most 4095 reactions");
  DevMechBlob blob;
  DevMechLayout& L = blob.lay;
  L.K = K;
  L.R = R;

  // Device reaction order: Troe, Lindemann, three-body, reversible, irreversible, so that
  // the 64-reaction chunks a wave evaluates together mostly take the same branch. The
  // order is stable within a group; results do not depend on it.
  std::vector<const Reaction*> rx_of(R);
  {
    std::vector<int> idx(R);
    for (int r = 0; r < R; ++r) idx[r] = r;
    auto group = [&](int r) {
      const Reaction& x = m.reactions[r];
      if (x.kind == kTroe) return 0;
      if (x.kind == kLindemann) return 1;
      if (x.kind == kThreeBody) return 2;
      return x.reversible ? 3 : 4;
    };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return group(a) < group(b); });
    for (int d = 0; d < R; ++d) rx_of[d] = &m.reactions[idx[d]];
  }

  // coefficient table: entry 0 is 0.0
  std::vector<double> coef{0.0};

  // net stoichiometry per reaction
  std::vector<std::map<int, double>> net(R);
  int n_pdep = 0, n_eff = 0;
  for (int r = 0; r < R; ++r) {
    const Reaction& rx = *rx_of[r];
    if (!rx.plog.empty())
      throw std::runtime_error("PLOG rates run in the block-per-cell layout: " + rx.equation);
    if (static_cast<int>(rx.reactants.size()) > kMaxSideSpecies ||
        static_cast<int>(rx.products.size()) > kMaxSideSpecies)
      throw std::runtime_error("more than 3 species on a reaction side: " + rx.equation);
    for (auto& kv : rx.reactants) {
      if (kv.second != std::floor(kv.second) || kv.second < 1 || kv.second > 3)
        throw std::runtime_error("non-integer stoichiometry unsupported: " + rx.equation);
      net[r][kv.first] -= kv.second;
    }
    for (auto& kv : rx.products) {
      if (kv.second != std::floor(kv.second) || kv.second < 1 || kv.second > 3)
        throw std::runtime_error("non-integer stoichiometry unsupported: " + rx.equation);
      net[r][kv.first] += kv.second;
    }
    if (rx.kind != kElementary) {
      ++n_pdep;
      n_eff += static_cast<int>(rx.eff.size());
    }
  }
  L.n_pdep = n_pdep;
  L.n_eff = n_eff;
  int max_eff = 0;
  for (int r = 0; r < R; ++r)
    if (rx_of[r]->kind != kElementary) max_eff = std::max(max_eff, static_cast<int>(rx_of[r]->eff.size()));
  L.n_eff_pad = align_to(std::max(max_eff, 1), 4);

  // balanced species gather over the nonzeros of the net stoichiometric matrix
  struct Nz {
    int k, r, c;
  };
  std::vector<Nz> nz;
  for (int k = 0; k < K; ++k)
    for (int r = 0; r < R; ++r) {
      auto it = net[r].find(k);
      if (it != net[r].end() && it->second != 0.0) {
        const int nu = static_cast<int>(it->second);
        if (nu < -8 || nu > 7) throw std::runtime_error("net stoichiometric coefficient out of range");
        nz.push_back({k, r, nu});
      }
    }
  const int NNZ = static_cast<int>(nz.size());
  const int per_lane = std::max(1, (NNZ + 63) / 64);
  L.n_gather_lane = per_lane;
  std::vector<uint32_t> gather(static_cast<size_t>(per_lane) * 64, 0u);
  std::vector<int> part_species;  // species of each partial slot
  for (int lane = 0; lane < 64; ++lane) {
    int b = lane * per_lane, e = std::min(NNZ, b + per_lane);
    for (int t = b; t < e; ++t) {
      bool flush = (t + 1 == e) || (nz[t + 1].k != nz[t].k);
      uint32_t part = 0;
      if (flush) {
        part = static_cast<uint32_t>(part_species.size());
        part_species.push_back(nz[t].k);
      }
      uint32_t w = static_cast<uint32_t>(nz[t].r) | ((static_cast<uint32_t>(nz[t].c) & 15u) << 12) |
                   (flush ? (1u << 16) : 0u) | (part << 17);
      gather[static_cast<size_t>(t - b) * 64 + lane] = w;
    }
    // padding entries: r=0, nu 0, no flush -> contribute nothing
  }
  L.n_partials = static_cast<int>(part_species.size());
  if (L.n_partials > 32767) throw std::runtime_error("too many gather partials");
  std::vector<uint16_t> part_ptr(K + 1, 0);
  {
    // partial slots of a species are contiguous because chunks are contiguous in species order
    int s = 0;
    for (int k = 0; k < K; ++k) {
      part_ptr[k] = static_cast<uint16_t>(s);
      while (s < L.n_partials && part_species[s] == k) ++s;
    }
    part_ptr[K] = static_cast<uint16_t>(s);
    if (s != L.n_partials) throw std::runtime_error("gather partials not contiguous");
  }

  // Jacobian slots: per reaction reactant slots, product slots (reversible), efficiency slots
  std::vector<uint16_t> slot_base(R + 1, 0);
  std::vector<int> slot_species;
  for (int r = 0; r < R; ++r) {
    const Reaction& rx = *rx_of[r];
    slot_base[r] = static_cast<uint16_t>(slot_species.size());
    for (auto& kv : rx.reactants) slot_species.push_back(kv.first);
    if (rx.reversible)
      for (auto& kv : rx.products) slot_species.push_back(kv.first);
    if (rx.kind != kElementary)
      for (auto& kv : rx.eff) slot_species.push_back(kv.first);
  }
  slot_base[R] = static_cast<uint16_t>(slot_species.size());
  L.n_slots = static_cast<int>(slot_species.size());
  if (L.n_slots > 65535) throw std::runtime_error("too many Jacobian slots");
  L.n_coef = static_cast<int>(coef.size());

  // Jacobian terms grouped by cell (col i, row): rows 0..K-2 species, row K = temperature
  std::map<int, std::vector<uint64_t>> cells;  // cell -> terms (src | coef<<16)
  for (int r = 0; r < R; ++r) {
    for (int s = slot_base[r]; s < slot_base[r + 1]; ++s) {
      int i = slot_species[s];
      for (auto& kv : net[r]) {
        int k = kv.first;
        if (kv.second == 0.0 || k == K - 1) continue;
        int cell = i * 64 + k;
        cells[cell].push_back(static_cast<uint64_t>(s) | (static_cast<uint64_t>(coef_index(coef, kv.second)) << 16));
      }
      int cell = i * 64 + K;
      cells[cell].push_back(static_cast<uint64_t>(s) | (static_cast<uint64_t>(16 + r) << 16));
    }
  }
  L.n_coef = static_cast<int>(coef.size());
  // greedy balanced assignment of cells to lanes
  std::vector<std::pair<int, int>> order;  // (terms, cell)
  for (auto& kv : cells) order.emplace_back(static_cast<int>(kv.second.size()), kv.first);
  std::sort(order.begin(), order.end(), [](auto& a, auto& b) { return a.first > b.first || (a.first == b.first && a.second < b.second); });
  std::vector<std::vector<int>> lane_cells(64);
  std::vector<int> load(64, 0);
  for (auto& oc : order) {
    int best = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
    lane_cells[best].push_back(oc.second);
    load[best] += oc.first;
  }
  int maxload = *std::max_element(load.begin(), load.end());
  L.n_jterm_lane = std::max(1, maxload);
  const uint64_t dummy = 4095ull | (1ull << 48);  // cell (63,63) unused, coef 0 -> writes 0
  std::vector<uint64_t> jterm(static_cast<size_t>(L.n_jterm_lane) * 64, dummy);
  for (int lane = 0; lane < 64; ++lane) {
    std::sort(lane_cells[lane].begin(), lane_cells[lane].end());
    int t = 0;
    for (int cell : lane_cells[lane]) {
      auto& v = cells[cell];
      for (size_t j = 0; j < v.size(); ++j) {
        uint64_t w = static_cast<uint64_t>(cell) | (v[j] << 16) | ((j + 1 == v.size()) ? (1ull << 48) : 0ull);
        jterm[static_cast<size_t>(t) * 64 + lane] = w;
        ++t;
      }
    }
  }

  // ---- layout ----
  int off = 0;
  L.off_species = off;
  off += K * kSpStride * 8;
  off = align_to(off, 16);
  L.off_hdr = off;
  off += R * static_cast<int>(sizeof(RxnHdr));
  L.off_mol = off;
  off += R * static_cast<int>(sizeof(RxnMol));
  L.off_arr = off;
  off += R * 3 * 8;
  L.off_pdep = off;
  off += n_pdep * static_cast<int>(sizeof(PdepRec));
  L.off_eff_val = off;
  off += n_eff * 8;
  L.off_eff_sp = off;
  off += align_to(std::max(n_eff, 1), 8);
  off = align_to(off, 16);
  L.off_effp_val = off;
  off += n_pdep * L.n_eff_pad * 8;
  L.off_effp_sp = off;
  off += align_to(std::max(n_pdep * L.n_eff_pad, 1), 16);
  L.off_coef = off;
  off += 16 * 8;
  L.off_gather = off;
  off += per_lane * 64 * 4;
  L.off_part_ptr = off;
  off += align_to((K + 1) * 2, 16);
  off = align_to(off, 16);
  L.off_exptab = off;
  off += 32 * 8;
  L.lds_bytes = off;
  L.off_slot_base = off;
  off += align_to((R + 1) * 2, 16);
  L.off_jterm = off;
  off += L.n_jterm_lane * 64 * 8;
  L.total_bytes = off;

  blob.bytes.assign(off, 0);
  uint8_t* B = blob.bytes.data();
  double* sp = reinterpret_cast<double*>(B + L.off_species);
  for (int k = 0; k < K; ++k) {
    sp[k * kSpStride + 0] = 1.0 / m.W[k];
    sp[k * kSpStride + 1] = m.W[k];
    sp[k * kSpStride + 2] = m.tmid[k];
    for (int j = 0; j < 14; ++j) sp[k * kSpStride + 3 + j] = m.nasa[14 * k + j];
  }
  RxnHdr* hdr = reinterpret_cast<RxnHdr*>(B + L.off_hdr);
  RxnMol* mol = reinterpret_cast<RxnMol*>(B + L.off_mol);
  L.mol_ok = 1;
  double* arr = reinterpret_cast<double*>(B + L.off_arr);
  PdepRec* pd = reinterpret_cast<PdepRec*>(B + L.off_pdep);
  double* effv = reinterpret_cast<double*>(B + L.off_eff_val);
  uint8_t* effs = B + L.off_eff_sp;
  int ip = 0, ie = 0;
  for (int r = 0; r < R; ++r) {
    const Reaction& rx = *rx_of[r];
    RxnHdr h;
    std::memset(&h, 0, sizeof(h));
    for (int s = 0; s < 6; ++s) h.sp[s] = kDummySpecies;  // nu = 0: contributes C^0 = 1, 0*g
    for (size_t s = 0; s < rx.reactants.size(); ++s) {
      h.sp[s] = static_cast<uint8_t>(rx.reactants[s].first);
      h.nu[s] = static_cast<uint8_t>(rx.reactants[s].second);
    }
    for (size_t s = 0; s < rx.products.size(); ++s) {
      h.sp[3 + s] = static_cast<uint8_t>(rx.products[s].first);
      h.nu[3 + s] = static_cast<uint8_t>(rx.products[s].second);
    }
    {
      RxnMol mo;
      std::memset(&mo, 0, sizeof(mo));
      for (int q = 0; q < 6; ++q) mo.m[q] = kDummySpecies;
      int nr = 0, np = 0;
      for (auto& kv : rx.reactants)
        for (int c = 0; c < static_cast<int>(kv.second); ++c) {
          if (nr < 3) mo.m[nr] = static_cast<uint8_t>(kv.first);
          ++nr;
        }
      for (auto& kv : rx.products)
        for (int c = 0; c < static_cast<int>(kv.second); ++c) {
          if (np < 3) mo.m[3 + np] = static_cast<uint8_t>(kv.first);
          ++np;
        }
      if (nr > 3 || np > 3) L.mol_ok = 0;
      mo.dnu = static_cast<int8_t>(np - nr);
      mol[r] = mo;
    }
    h.kind = static_cast<uint8_t>(rx.kind);
    bool constk = (rx.b == 0.0 && rx.theta == 0.0);
    h.flags = (rx.reversible ? 1 : 0) | (constk ? 2 : 0) | (rx.troe_has_t2 ? 4 : 0);
    arr[3 * r + 0] = constk ? rx.A : rx.lnA;
    arr[3 * r + 1] = rx.b;
    arr[3 * r + 2] = rx.theta;
    if (rx.kind != kElementary) {
      h.pidx = static_cast<uint16_t>(ip);
      PdepRec& P = pd[ip];
      P.eff_base = rx.eff_base;
      P.eff_begin = ie;
      for (auto& kv : rx.eff) {
        effv[ie] = kv.second;
        effs[ie] = static_cast<uint8_t>(kv.first);
        ++ie;
      }
      P.eff_end = ie;
      double* pv = reinterpret_cast<double*>(B + L.off_effp_val) + ip * L.n_eff_pad;
      uint8_t* ps = B + L.off_effp_sp + ip * L.n_eff_pad;
      for (int j = 0; j < L.n_eff_pad; ++j) {
        const bool real = j < static_cast<int>(rx.eff.size());
        pv[j] = real ? rx.eff[j].second : 0.0;
        ps[j] = real ? static_cast<uint8_t>(rx.eff[j].first) : kDummySpecies;
      }
      bool constk0 = (rx.b0 == 0.0 && rx.theta0 == 0.0);
      if (rx.kind >= kLindemann && constk0) h.flags |= 8;
      P.lnA0 = (rx.kind >= kLindemann && constk0) ? rx.A0 : rx.lnA0;
      P.b0 = rx.b0;
      P.theta0 = rx.theta0;
      P.troe_a = rx.troe[0];
      troe_device_coefs(rx, P.troe_iT3, P.troe_iT1, P.troe_T2);
      ++ip;
    }
    hdr[r] = h;
  }
  double* ctab = reinterpret_cast<double*>(B + L.off_coef);
  for (size_t i = 0; i < coef.size(); ++i) ctab[i] = coef[i];
  std::memcpy(B + L.off_gather, gather.data(), gather.size() * 4);
  std::memcpy(B + L.off_part_ptr, part_ptr.data(), part_ptr.size() * 2);
  std::memcpy(B + L.off_exptab, kExp2Tab32, sizeof(kExp2Tab32));
  std::memcpy(B + L.off_slot_base, slot_base.data(), slot_base.size() * 2);
  std::memcpy(B + L.off_jterm, jterm.data(), jterm.size() * 8);
  return blob;
}

DevBlkBlob build_block_tables(const Mechanism& m, int work_lds_bytes) {
  const int K = m.K(), R = m.R();
  if (K < 2 || K > 255) throw std::runtime_error("block-per-cell path supports 2..255 species");
  if (R > 65535) throw std::runtime_error("block-per-cell path supports at most 65535 reactions");
  DevBlkBlob blob;
  BlkLayout& L = blob.lay;
  L.K = K;
  L.R = R;
  // device order as on the wave path: Troe, Lindemann, three-body, reversible, irreversible
  std::vector<const Reaction*> rx_of(R);
  {
    std::vector<int> idx(R);
    for (int r = 0; r < R; ++r) idx[r] = r;
    auto group = [&](int r) {
      const Reaction& x = m.reactions[r];
      if (x.kind == kTroe) return 0;
      if (x.kind == kLindemann) return 1;
      if (x.kind == kThreeBody) return 2;
      return x.reversible ? 3 : 4;
    };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return group(a) < group(b); });
    for (int d = 0; d < R; ++d) rx_of[d] = &m.reactions[idx[d]];
  }
  auto integer_stoich = [](const std::vector<std::pair<int, double>>& side) {
    for (auto& kv : side)
      if (kv.second != std::floor(kv.second) || kv.second < 1 || kv.second > 3) return false;
    return true;
  };
  // distinct net coefficients, indexed in the species lists
  std::vector<double> coefs;
  auto coef_index = [&](double c) {
    for (size_t i = 0; i < coefs.size(); ++i)
      if (coefs[i] == c) return static_cast<uint32_t>(i);
    if (coefs.size() >= 65535) throw std::runtime_error("too many distinct stoichiometric coefficients");
    coefs.push_back(c);
    return static_cast<uint32_t>(coefs.size() - 1);
  };
  std::vector<std::map<int, double>> net(R);
  std::vector<bool> lumped(R, false);
  int n_pdep = 0, n_eff = 0;
  std::vector<int> slot_sp;
  std::vector<uint16_t> slot_begin(R), slot_count(R);
  for (int r = 0; r < R; ++r) {
    const Reaction& rx = *rx_of[r];
    if (static_cast<int>(rx.reactants.size()) > kBlkSide)
      throw std::runtime_error("more than 4 reactant species: " + rx.equation);
    if (!integer_stoich(rx.reactants)) throw std::runtime_error("non-integer reactant stoichiometry unsupported: " + rx.equation);
    const bool prod_ok = static_cast<int>(rx.products.size()) <= kBlkSide && integer_stoich(rx.products);
    if (!prod_ok) {
      if (rx.reversible)
        throw std::runtime_error("a reversible reaction needs at most 4 products with integer coefficients: " +
                                 rx.equation);
      for (auto& kv : rx.products)
        if (!(kv.second > 0)) throw std::runtime_error("non-positive product coefficient: " + rx.equation);
      lumped[r] = true;
    }
    for (auto& kv : rx.reactants) net[r][kv.first] -= kv.second;
    for (auto& kv : rx.products) net[r][kv.first] += kv.second;
    slot_begin[r] = static_cast<uint16_t>(slot_sp.size());
    for (auto& kv : rx.reactants) slot_sp.push_back(kv.first);
    if (rx.reversible)
      for (auto& kv : rx.products) slot_sp.push_back(kv.first);
    if (rx.kind != kElementary) {
      ++n_pdep;
      n_eff += static_cast<int>(rx.eff.size());
      for (auto& kv : rx.eff) slot_sp.push_back(kv.first);
    }
    if (slot_sp.size() > 65535) throw std::runtime_error("too many Jacobian slots");
    slot_count[r] = static_cast<uint16_t>(slot_sp.size() - slot_begin[r]);
  }
  L.n_pdep = n_pdep;
  L.n_eff = n_eff;
  L.n_slots = static_cast<int>(slot_sp.size());
  // per-species reaction lists (net coefficient != 0), ascending device reaction index
  std::vector<uint16_t> csr_ptr(K + 1, 0);
  std::vector<uint32_t> csr;
  for (int k = 0; k < K; ++k) {
    csr_ptr[k] = static_cast<uint16_t>(csr.size());
    for (int r = 0; r < R; ++r) {
      auto it = net[r].find(k);
      if (it == net[r].end() || it->second == 0) continue;
      csr.push_back(static_cast<uint32_t>(r) | (coef_index(it->second) << 16));
    }
    if (csr.size() > 65535) throw std::runtime_error("too many species-reaction pairs");
  }
  csr_ptr[K] = static_cast<uint16_t>(csr.size());
  L.n_csr = static_cast<int>(csr.size());
  // products of lumped reactions (their enthalpy and mole change)
  std::vector<uint16_t> lump_ptr(R + 1, 0);
  std::vector<uint32_t> lump;
  for (int r = 0; r < R; ++r) {
    lump_ptr[r] = static_cast<uint16_t>(lump.size());
    if (lumped[r])
      for (auto& kv : rx_of[r]->products) lump.push_back(static_cast<uint32_t>(kv.first) | (coef_index(kv.second) << 16));
    if (lump.size() > 65535) throw std::runtime_error("too many lumped products");
  }
  lump_ptr[R] = static_cast<uint16_t>(lump.size());
  L.n_lump = static_cast<int>(lump.size());
  L.n_coef = static_cast<int>(coefs.size());
  // PLOG rates: device reaction, pressure levels
  std::vector<uint16_t> plog_rxn, plog_ptr(1, 0);
  std::vector<double> plog_lv;
  for (int r = 0; r < R; ++r) {
    const Reaction& rx = *rx_of[r];
    if (rx.plog.empty()) continue;
    plog_rxn.push_back(static_cast<uint16_t>(r));
    for (const PlogLevel& lv : rx.plog) {
      plog_lv.push_back(lv.lnP);
      plog_lv.push_back(lv.lnA);
      plog_lv.push_back(lv.b);
      plog_lv.push_back(lv.theta);
    }
    if (plog_lv.size() / 4 > 65535) throw std::runtime_error("too many PLOG levels");
    plog_ptr.push_back(static_cast<uint16_t>(plog_lv.size() / 4));
  }
  L.n_plog = static_cast<int>(plog_rxn.size());
  L.n_plog_lv = static_cast<int>(plog_lv.size() / 4);
  // per-species Jacobian slots (the temperature row's column sums)
  std::vector<uint32_t> col_ptr(K + 1, 0);
  std::vector<uint32_t> col;
  for (int k = 0; k < K; ++k) {
    col_ptr[k] = static_cast<uint32_t>(col.size());
    for (int r = 0; r < R; ++r)
      for (int q = 0; q < slot_count[r]; ++q)
        if (slot_sp[slot_begin[r] + q] == k)
          col.push_back(static_cast<uint32_t>(slot_begin[r] + q) | (static_cast<uint32_t>(r) << 16));
  }
  col_ptr[K] = static_cast<uint32_t>(col.size());
  L.n_col = static_cast<int>(col.size());

  // tables in staging order; the LDS prefix ends at the last table that fits
  int off = 0;
  std::vector<int> ends;
  auto place = [&](int32_t& field, int bytes) {
    field = off;
    off = align_to(off + std::max(bytes, 1), 16);
    ends.push_back(off);
  };
  place(L.off_exptab, 32 * 8);
  place(L.off_arr, R * 3 * 8);
  const int min_staged = off;
  place(L.off_species, K * kSpStride * 8);
  place(L.off_rxn, R * static_cast<int>(sizeof(BlkRxn)));
  place(L.off_pdep, n_pdep * static_cast<int>(sizeof(PdepRec)));
  place(L.off_eff_val, n_eff * 8);
  place(L.off_eff_sp, n_eff);
  place(L.off_csr_ptr, (K + 1) * 2);
  place(L.off_coef, L.n_coef * 8);
  place(L.off_csr, L.n_csr * 4);
  const int budget = kBlkLdsMax - work_lds_bytes;
  if (min_staged > budget)
    throw std::runtime_error("the Arrhenius table and the work area exceed the LDS of a CU");
  L.lds_bytes = min_staged;
  for (int e : ends)
    if (e <= budget) L.lds_bytes = std::max(L.lds_bytes, e);
  place(L.off_slot_sp, L.n_slots);
  place(L.off_col_ptr, (K + 1) * 4);
  place(L.off_col, L.n_col * 4);
  place(L.off_plog_rxn, L.n_plog * 2);
  place(L.off_plog_ptr, (L.n_plog + 1) * 2);
  place(L.off_plog_lv, L.n_plog_lv * 4 * 8);
  place(L.off_lump_ptr, (R + 1) * 2);
  place(L.off_lump, L.n_lump * 4);
  L.total_bytes = off;

  blob.bytes.assign(off, 0);
  uint8_t* B = blob.bytes.data();
  double* sp = reinterpret_cast<double*>(B + L.off_species);
  for (int k = 0; k < K; ++k) {
    sp[k * kSpStride + 0] = 1.0 / m.W[k];
    sp[k * kSpStride + 1] = m.W[k];
    sp[k * kSpStride + 2] = m.tmid[k];
    for (int j = 0; j < 14; ++j) sp[k * kSpStride + 3 + j] = m.nasa[14 * k + j];
  }
  BlkRxn* rxn = reinterpret_cast<BlkRxn*>(B + L.off_rxn);
  double* arr = reinterpret_cast<double*>(B + L.off_arr);
  PdepRec* pd = reinterpret_cast<PdepRec*>(B + L.off_pdep);
  double* effv = reinterpret_cast<double*>(B + L.off_eff_val);
  uint8_t* effs = B + L.off_eff_sp;
  int ip = 0, ie = 0;
  for (int r = 0; r < R; ++r) {
    const Reaction& rx = *rx_of[r];
    BlkRxn h;
    std::memset(&h, 0, sizeof(h));
    for (int q = 0; q < 8; ++q) h.sp[q] = static_cast<uint8_t>(K);
    for (size_t q = 0; q < rx.reactants.size(); ++q) {
      h.sp[q] = static_cast<uint8_t>(rx.reactants[q].first);
      h.nu[q] = static_cast<uint8_t>(rx.reactants[q].second);
    }
    if (!lumped[r])
      for (size_t q = 0; q < rx.products.size(); ++q) {
        h.sp[kBlkSide + q] = static_cast<uint8_t>(rx.products[q].first);
        h.nu[kBlkSide + q] = static_cast<uint8_t>(rx.products[q].second);
      }
    h.kind = static_cast<uint8_t>(rx.kind);
    // a PLOG rate's Arrhenius parameters change with the cell's pressure: never a constant
    const bool constk = rx.plog.empty() && (rx.b == 0.0 && rx.theta == 0.0);
    h.flags = (rx.reversible ? 1 : 0) | (constk ? 2 : 0) | (rx.troe_has_t2 ? 4 : 0) | (lumped[r] ? kBlkLumped : 0);
    h.pidx = 0xFFFF;
    h.slot = slot_begin[r];
    h.nslot = slot_count[r];
    arr[3 * r + 0] = constk ? rx.A : rx.lnA;
    arr[3 * r + 1] = rx.b;
    arr[3 * r + 2] = rx.theta;
    if (rx.kind != kElementary) {
      h.pidx = static_cast<uint16_t>(ip);
      PdepRec& P = pd[ip];
      P.eff_base = rx.eff_base;
      P.eff_begin = ie;
      for (auto& kv : rx.eff) {
        effv[ie] = kv.second;
        effs[ie] = static_cast<uint8_t>(kv.first);
        ++ie;
      }
      P.eff_end = ie;
      const bool constk0 = (rx.b0 == 0.0 && rx.theta0 == 0.0);
      if (rx.kind >= kLindemann && constk0) h.flags |= 8;
      P.lnA0 = (rx.kind >= kLindemann && constk0) ? rx.A0 : rx.lnA0;
      P.b0 = rx.b0;
      P.theta0 = rx.theta0;
      P.troe_a = rx.troe[0];
      troe_device_coefs(rx, P.troe_iT3, P.troe_iT1, P.troe_T2);
      ++ip;
    }
    rxn[r] = h;
  }
  std::memcpy(B + L.off_exptab, kExp2Tab32, sizeof(kExp2Tab32));
  std::memcpy(B + L.off_csr_ptr, csr_ptr.data(), csr_ptr.size() * 2);
  if (!coefs.empty()) std::memcpy(B + L.off_coef, coefs.data(), coefs.size() * 8);
  if (!csr.empty()) std::memcpy(B + L.off_csr, csr.data(), csr.size() * 4);
  for (int q = 0; q < L.n_slots; ++q) B[L.off_slot_sp + q] = static_cast<uint8_t>(slot_sp[q]);
  std::memcpy(B + L.off_col_ptr, col_ptr.data(), col_ptr.size() * 4);
  if (!col.empty()) std::memcpy(B + L.off_col, col.data(), col.size() * 4);
  if (!plog_rxn.empty()) std::memcpy(B + L.off_plog_rxn, plog_rxn.data(), plog_rxn.size() * 2);
  std::memcpy(B + L.off_plog_ptr, plog_ptr.data(), plog_ptr.size() * 2);
  if (!plog_lv.empty()) std::memcpy(B + L.off_plog_lv, plog_lv.data(), plog_lv.size() * 8);
  std::memcpy(B + L.off_lump_ptr, lump_ptr.data(), lump_ptr.size() * 2);
  if (!lump.empty()) std::memcpy(B + L.off_lump, lump.data(), lump.size() * 4);
  return blob;
}

}  // namespace hpo
