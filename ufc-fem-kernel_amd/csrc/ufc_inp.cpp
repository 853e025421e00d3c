// ufc_inp.cpp -- reads a static C3D8 step from an Abaqus-style input deck into
// the arrays of the C-ABI (ufc_ap_model).  Host code only.
//
// The reference reads decks through AP_Input_Domain_ParseKeyword
// (L6_AP/Input/Parser/AP_Inp_Mgr.f90:99-125) and its keyword table
// (AP_Inp_KWBrg.f90); this reader covers the keywords of the static
// stiffness path and reports everything else it cannot represent as
// UFC_STATUS_UNSUPPORTED instead of guessing.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "ufc_hip.h"

namespace {

struct InpError {
  int code;
  std::string msg;
  int line = 0;
};

[[noreturn]] void raise(int code, int line, const std::string& msg) {
  throw InpError{code, "line " + std::to_string(line) + ": " + msg, line};
}

std::string upper(std::string s) {
  for (auto& ch : s) ch = char(std::toupper(static_cast<unsigned char>(ch)));
  return s;
}

std::string trim(const std::string& s) {
  size_t b = 0, e = s.size();
  while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
  while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
  return s.substr(b, e - b);
}

std::vector<std::string> split_commas(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : s) {
    if (ch == ',') {
      out.push_back(trim(cur));
      cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  out.push_back(trim(cur));
  return out;
}

struct Keyword {
  std::string name;                           // upper case, single spaces
  std::map<std::string, std::string> params;  // KEY -> VALUE (upper-case key)
  int line = 0;
};

Keyword parse_keyword(const std::string& text, int line) {
  Keyword k;
  k.line = line;
  auto parts = split_commas(text.substr(1));
  std::string nm;
  for (char ch : upper(parts[0])) {
    if (std::isspace(static_cast<unsigned char>(ch))) {
      if (!nm.empty() && nm.back() != ' ') nm.push_back(' ');
    } else {
      nm.push_back(ch);
    }
  }
  k.name = trim(nm);
  for (size_t i = 1; i < parts.size(); ++i) {
    if (parts[i].empty()) continue;
    const size_t eq = parts[i].find('=');
    if (eq == std::string::npos)
      k.params[upper(trim(parts[i]))] = "";
    else
      k.params[upper(trim(parts[i].substr(0, eq)))] = trim(parts[i].substr(eq + 1));
  }
  return k;
}

bool is_int(const std::string& s) {
  if (s.empty()) return false;
  size_t i = (s[0] == '-' || s[0] == '+') ? 1 : 0;
  if (i == s.size()) return false;
  for (; i < s.size(); ++i)
    if (!std::isdigit(static_cast<unsigned char>(s[i]))) return false;
  return true;
}

long to_int(const std::string& s, int line) {
  if (!is_int(s)) raise(UFC_STATUS_INVALID, line, "expected an integer, got '" + s + "'");
  return std::stol(s);
}

double to_real(const std::string& s, int line) {
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (s.empty() || end == s.c_str() || *end != '\0')
    raise(UFC_STATUS_INVALID, line, "expected a number, got '" + s + "'");
  return v;
}

// Face labels S1..S6 (and P1..P6 of *DLOAD) name column k of
// PH_ELEM_C3D8_FACE_NODES (PH_Elem_C3D8.f90:120-121) directly, as the
// reference's *SURFACE mapping does (MD_KW_Mapper.f90:2137-2155): "S4" is the
// face (4,3,7,8), which Abaqus calls S5.  Returned 0-based.
int face_index(const std::string& label, int line) {
  const std::string u = upper(label);
  if (u.size() == 2 && (u[0] == 'S' || u[0] == 'P') && u[1] >= '1' && u[1] <= '6')
    return u[1] - '1';
  raise(UFC_STATUS_UNSUPPORTED, line, "face label '" + label + "' (C3D8 faces are S1..S6 / P1..P6)");
}

struct Model {
  // nodes and elements in order of appearance
  std::vector<int32_t> node_label, elem_label;
  std::vector<double> coords;
  std::vector<int32_t> conn_label;  // labels, resolved at the end
  std::map<int32_t, int32_t> node_of, elem_of;
  std::map<std::string, std::vector<int32_t>> nset, elset;  // labels
  std::map<std::string, std::vector<std::pair<std::string, int>>> surface;  // (elset/elem, face)
  std::map<std::string, std::pair<double, double>> material;  // E, nu
  std::vector<std::pair<std::string, std::string>> sections;  // elset, material
  std::map<int32_t, double> bc;                                // eq -> value
  std::vector<std::pair<int32_t, double>> cload;               // eq, value
  std::vector<std::tuple<int32_t, int, double>> faces;         // elem label, face, p
  // results
  std::vector<int32_t> conn, bc_eq, cload_eq, face_elem, face_id;
  std::vector<double> young, bc_val, cload_val, face_p;
  double nu = 0.0;
  int variant = 0;
};

class Reader {
 public:
  explicit Reader(Model& m) : m_(m) {}

  void run(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw InpError{UFC_STATUS_INVALID, "cannot open '" + path + "'"};
    std::vector<std::pair<int, std::string>> lines;
    std::string s;
    int ln = 0;
    while (std::getline(in, s)) {
      ++ln;
      if (!s.empty() && s.back() == '\r') s.pop_back();
      const std::string t = trim(s);
      if (t.empty() || t.rfind("**", 0) == 0) continue;
      lines.emplace_back(ln, t);
    }
    Keyword kw;
    bool have_kw = false;
    std::string pending;  // element data continued over lines
    int pending_line = 0;
    for (auto& [line, text] : lines) {
      if (text[0] == '*') {
        if (!pending.empty()) raise(UFC_STATUS_INVALID, pending_line, "incomplete element line");
        kw = parse_keyword(text, line);
        have_kw = true;
        begin(kw);
        continue;
      }
      if (!have_kw) raise(UFC_STATUS_INVALID, line, "data before the first keyword");
      if (kw.name == "ELEMENT") {
        // an element's data may continue on the next line after a comma
        pending += text;
        if (pending_line == 0) pending_line = line;
        if (text.back() == ',') continue;
        data(kw, pending, pending_line);
        pending.clear();
        pending_line = 0;
      } else {
        data(kw, text, line);
      }
    }
    if (!pending.empty()) raise(UFC_STATUS_INVALID, pending_line, "incomplete element line");
    finish();
  }

 private:
  Model& m_;
  std::string part_;               // part being defined
  int n_parts_ = 0;
  std::map<std::string, std::string> instance_part_;
  std::string material_;           // material being defined
  int n_steps_ = 0;
  bool in_step_ = false;

  // "INST.NAME" -> "NAME" when INST is the instance of the deck's part
  std::string unqualify(const std::string& ref, int line) const {
    const size_t dot = ref.find('.');
    if (dot == std::string::npos) return ref;
    const std::string inst = upper(ref.substr(0, dot));
    if (!instance_part_.count(inst))
      raise(UFC_STATUS_INVALID, line, "unknown instance '" + ref.substr(0, dot) + "'");
    return ref.substr(dot + 1);
  }

  std::vector<int32_t> node_refs(const std::string& ref, int line) const {
    const std::string r = unqualify(ref, line);
    if (is_int(r)) return {int32_t(to_int(r, line))};
    auto it = m_.nset.find(upper(r));
    if (it == m_.nset.end()) raise(UFC_STATUS_INVALID, line, "unknown node set '" + r + "'");
    return it->second;
  }

  std::vector<int32_t> elem_refs(const std::string& ref, int line) const {
    const std::string r = unqualify(ref, line);
    if (is_int(r)) return {int32_t(to_int(r, line))};
    auto it = m_.elset.find(upper(r));
    if (it == m_.elset.end()) raise(UFC_STATUS_INVALID, line, "unknown element set '" + r + "'");
    return it->second;
  }

  static std::string param(const Keyword& k, const char* key) {
    auto it = k.params.find(key);
    return it == k.params.end() ? std::string() : it->second;
  }

  void begin(const Keyword& k) {
    const std::string& n = k.name;
    if (n == "PART") {
      if (++n_parts_ > 1) raise(UFC_STATUS_UNSUPPORTED, k.line, "more than one *PART");
      part_ = upper(param(k, "NAME"));
    } else if (n == "END PART") {
      part_.clear();
    } else if (n == "INSTANCE") {
      const std::string p = upper(param(k, "PART"));
      if (!instance_part_.empty())
        raise(UFC_STATUS_UNSUPPORTED, k.line, "more than one *INSTANCE");
      instance_part_[upper(param(k, "NAME"))] = p;
    } else if (n == "ELEMENT") {
      // C3D8 family -> PH_ELEM_C3D8_VARIANT_* (PH_Elem_C3D8.f90:110-116)
      static const std::map<std::string, int> kVariant = {
          {"C3D8", 1}, {"C3D8R", 2}, {"C3D8H", 3}, {"C3D8I", 4}, {"C3D8M", 5}, {"C3D8S", 6}};
      const auto it = kVariant.find(upper(param(k, "TYPE")));
      if (it == kVariant.end())
        raise(UFC_STATUS_UNSUPPORTED, k.line,
              "element type '" + param(k, "TYPE") + "' (this path reads the C3D8 family)");
      if (m_.variant != 0 && m_.variant != it->second)
        raise(UFC_STATUS_UNSUPPORTED, k.line, "several C3D8 variants in one model");
      m_.variant = it->second;
    } else if (n == "MATERIAL") {
      material_ = upper(param(k, "NAME"));
      m_.material[material_] = {0.0, -1.0};
    } else if (n == "SOLID SECTION") {
      m_.sections.emplace_back(upper(param(k, "ELSET")), upper(param(k, "MATERIAL")));
    } else if (n == "STEP") {
      if (++n_steps_ > 1) raise(UFC_STATUS_UNSUPPORTED, k.line, "more than one *STEP");
      const std::string nl = upper(param(k, "NLGEOM"));
      if (!nl.empty() && nl != "NO")
        raise(UFC_STATUS_UNSUPPORTED, k.line, "*STEP, NLGEOM (this path is small-strain)");
      in_step_ = true;
    } else if (n == "END STEP") {
      in_step_ = false;
    } else if (n == "STATIC") {
    } else if (n == "DYNAMIC" || n == "FREQUENCY" || n == "BUCKLE" || n == "HEAT TRANSFER") {
      raise(UFC_STATUS_UNSUPPORTED, k.line, "*" + n + " (this path is a static step)");
    } else if (n == "ELASTIC") {
      const std::string t = upper(param(k, "TYPE"));
      if (!t.empty() && t != "ISOTROPIC")
        raise(UFC_STATUS_UNSUPPORTED, k.line, "*ELASTIC, TYPE=" + t);
    } else if (n == "SURFACE") {
      const std::string t = upper(param(k, "TYPE"));
      if (!t.empty() && t != "ELEMENT")
        raise(UFC_STATUS_UNSUPPORTED, k.line, "*SURFACE, TYPE=" + t);
      m_.surface[upper(param(k, "NAME"))];
    } else if (n == "BOUNDARY" || n == "CLOAD" || n == "DLOAD" || n == "DSLOAD") {
      // the step is linear: AMPLITUDE ramps reach their full value at its end
      for (auto& kv : k.params) {
        const std::string v = upper(kv.second);
        if (kv.first == "OP" && (v == "MOD" || v == "NEW")) continue;
        if (kv.first == "AMPLITUDE") continue;
        if (n == "BOUNDARY" && kv.first == "TYPE" && v == "DISPLACEMENT") continue;
        raise(UFC_STATUS_UNSUPPORTED, k.line, "*" + n + " parameter " + kv.first);
      }
      // OP=NEW removes what earlier keywords of the same kind defined
      if (upper(param(k, "OP")) == "NEW") {
        if (n == "BOUNDARY") m_.bc.clear();
        if (n == "CLOAD") m_.cload.clear();
        if (n == "DLOAD" || n == "DSLOAD") m_.faces.clear();
      }
    } else if (n == "NSET") {
      m_.nset[upper(param(k, "NSET"))];
    } else if (n == "ELSET") {
      m_.elset[upper(param(k, "ELSET"))];
    }
  }

  void data(const Keyword& k, const std::string& text, int line) {
    const std::string& n = k.name;
    auto f = split_commas(text);
    while (!f.empty() && f.back().empty()) f.pop_back();
    if (n == "INSTANCE") {
      raise(UFC_STATUS_UNSUPPORTED, line, "*INSTANCE translation or rotation");
    } else if (n == "NODE") {
      if (f.size() < 3) raise(UFC_STATUS_INVALID, line, "node line needs a label and coordinates");
      const int32_t lab = int32_t(to_int(f[0], line));
      if (m_.node_of.count(lab)) raise(UFC_STATUS_INVALID, line, "node " + f[0] + " defined twice");
      m_.node_of[lab] = int32_t(m_.node_label.size());
      m_.node_label.push_back(lab);
      for (int d = 0; d < 3; ++d)
        m_.coords.push_back(size_t(d + 1) < f.size() ? to_real(f[size_t(d) + 1], line) : 0.0);
      const std::string ns = upper(param(k, "NSET"));
      if (!ns.empty()) m_.nset[ns].push_back(lab);
    } else if (n == "ELEMENT") {
      if (f.size() != 9) raise(UFC_STATUS_INVALID, line, "C3D8 line needs a label and 8 nodes");
      const int32_t lab = int32_t(to_int(f[0], line));
      if (m_.elem_of.count(lab))
        raise(UFC_STATUS_INVALID, line, "element " + f[0] + " defined twice");
      m_.elem_of[lab] = int32_t(m_.elem_label.size());
      m_.elem_label.push_back(lab);
      for (int a = 1; a <= 8; ++a) m_.conn_label.push_back(int32_t(to_int(f[size_t(a)], line)));
      const std::string es = upper(param(k, "ELSET"));
      if (!es.empty()) m_.elset[es].push_back(lab);
    } else if (n == "NSET" || n == "ELSET") {
      auto& set = n == "NSET" ? m_.nset[upper(param(k, "NSET"))]
                              : m_.elset[upper(param(k, "ELSET"))];
      if (k.params.count("GENERATE")) {
        if (f.size() < 2) raise(UFC_STATUS_INVALID, line, "GENERATE needs first, last[, step]");
        const long a = to_int(f[0], line), b = to_int(f[1], line);
        const long st = f.size() > 2 ? to_int(f[2], line) : 1;
        if (st <= 0 || b < a) raise(UFC_STATUS_INVALID, line, "bad GENERATE range");
        for (long v = a; v <= b; v += st) set.push_back(int32_t(v));
      } else {
        for (auto& t : f) {
          if (t.empty()) continue;
          const std::string r = unqualify(t, line);
          if (is_int(r)) {
            set.push_back(int32_t(to_int(r, line)));
          } else {
            auto& src = n == "NSET" ? m_.nset : m_.elset;
            auto it = src.find(upper(r));
            if (it == src.end()) raise(UFC_STATUS_INVALID, line, "unknown set '" + r + "'");
            set.insert(set.end(), it->second.begin(), it->second.end());
          }
        }
      }
    } else if (n == "SURFACE") {
      if (f.size() < 2) raise(UFC_STATUS_INVALID, line, "surface line needs elements and a face");
      m_.surface[upper(param(k, "NAME"))].emplace_back(f[0], face_index(f[1], line));
    } else if (n == "ELASTIC") {
      if (material_.empty()) raise(UFC_STATUS_INVALID, line, "*ELASTIC outside a *MATERIAL");
      if (f.size() < 2) raise(UFC_STATUS_INVALID, line, "*ELASTIC needs E, nu");
      auto& mat = m_.material[material_];
      if (mat.second < 0.0) mat = {to_real(f[0], line), to_real(f[1], line)};
    } else if (n == "BOUNDARY") {
      if (f.size() < 2) raise(UFC_STATUS_INVALID, line, "boundary line needs nodes and a dof");
      int d0 = 0, d1 = -1;
      double val = 0.0;
      const std::string t = upper(f[1]);
      if (t == "ENCASTRE" || t == "PINNED") {
        d0 = 1;
        d1 = 3;
      } else if (t == "XSYMM" || t == "YSYMM" || t == "ZSYMM") {
        d0 = d1 = 1 + (t[0] - 'X');
      } else {
        d0 = int(to_int(f[1], line));
        d1 = f.size() > 2 && !f[2].empty() ? int(to_int(f[2], line)) : d0;
        if (f.size() > 3) val = to_real(f[3], line);
      }
      if (d0 < 1 || d1 < d0) raise(UFC_STATUS_INVALID, line, "bad dof range");
      // degrees of freedom 4-6 (rotations) do not exist on solid nodes
      for (int32_t lab : node_refs(f[0], line))
        for (int d = d0; d <= std::min(d1, 3); ++d) m_.bc[node_eq(lab, d, line)] = val;
    } else if (n == "CLOAD") {
      if (f.size() < 3) raise(UFC_STATUS_INVALID, line, "*CLOAD line needs nodes, dof, value");
      const int d = int(to_int(f[1], line));
      const double v = to_real(f[2], line);
      if (d < 1 || d > 6) raise(UFC_STATUS_INVALID, line, "bad dof");
      if (d > 3) raise(UFC_STATUS_UNSUPPORTED, line, "moment on a solid node");
      for (int32_t lab : node_refs(f[0], line)) m_.cload.emplace_back(node_eq(lab, d, line), v);
    } else if (n == "DLOAD") {
      if (f.size() < 3) raise(UFC_STATUS_INVALID, line, "*DLOAD line needs elements, Pn, value");
      const int face = face_index(f[1], line);
      const double p = to_real(f[2], line);
      for (int32_t lab : elem_refs(f[0], line)) m_.faces.emplace_back(lab, face, p);
    } else if (n == "DSLOAD") {
      if (f.size() < 3) raise(UFC_STATUS_INVALID, line, "*DSLOAD line needs a surface, P, value");
      if (upper(f[1]) != "P") raise(UFC_STATUS_UNSUPPORTED, line, "*DSLOAD type " + f[1]);
      const double p = to_real(f[2], line);
      auto it = m_.surface.find(upper(unqualify(f[0], line)));
      if (it == m_.surface.end()) raise(UFC_STATUS_INVALID, line, "unknown surface '" + f[0] + "'");
      for (auto& [ref, face] : it->second)
        for (int32_t lab : elem_refs(ref, line)) m_.faces.emplace_back(lab, face, p);
    }
    // other keywords' data lines are not part of this path
  }

  int32_t node_eq(int32_t lab, int dof, int line) const {
    auto it = m_.node_of.find(lab);
    if (it == m_.node_of.end()) raise(UFC_STATUS_INVALID, line, "unknown node " + std::to_string(lab));
    return 3 * it->second + (dof - 1);
  }

  void finish() {
    const size_t ne = m_.elem_label.size();
    if (m_.node_label.empty() || ne == 0)
      throw InpError{UFC_STATUS_INVALID, "the deck defines no C3D8-family elements"};
    m_.conn.resize(m_.conn_label.size());
    for (size_t i = 0; i < m_.conn_label.size(); ++i) {
      auto it = m_.node_of.find(m_.conn_label[i]);
      if (it == m_.node_of.end())
        throw InpError{UFC_STATUS_INVALID, "element " + std::to_string(m_.elem_label[i / 8]) +
                                               " uses undefined node " +
                                               std::to_string(m_.conn_label[i])};
      m_.conn[i] = it->second;
    }
    // material per element: sections, or the only material of the deck
    std::vector<int> mat_of(ne, -1);
    std::vector<std::string> mats;
    for (auto& kv : m_.material) {
      if (kv.second.second < 0.0)
        throw InpError{UFC_STATUS_INVALID, "material " + kv.first + " has no *ELASTIC"};
      mats.push_back(kv.first);
    }
    auto mat_index = [&](const std::string& nm) {
      auto it = std::find(mats.begin(), mats.end(), nm);
      if (it == mats.end()) throw InpError{UFC_STATUS_INVALID, "unknown material " + nm};
      return int(it - mats.begin());
    };
    if (m_.sections.empty()) {
      if (mats.size() != 1)
        throw InpError{UFC_STATUS_INVALID, "no *SOLID SECTION and not exactly one material"};
      std::fill(mat_of.begin(), mat_of.end(), 0);
    } else {
      for (auto& [es, mat] : m_.sections) {
        auto it = m_.elset.find(es);
        if (it == m_.elset.end()) throw InpError{UFC_STATUS_INVALID, "unknown element set " + es};
        const int mi = mat_index(mat);
        for (int32_t lab : it->second) {
          auto e = m_.elem_of.find(lab);
          if (e == m_.elem_of.end())
            throw InpError{UFC_STATUS_INVALID, "section on undefined element " + std::to_string(lab)};
          mat_of[size_t(e->second)] = mi;
        }
      }
    }
    m_.young.resize(ne);
    std::set<double> nus;
    for (size_t e = 0; e < ne; ++e) {
      if (mat_of[e] < 0)
        throw InpError{UFC_STATUS_INVALID,
                       "element " + std::to_string(m_.elem_label[e]) + " has no section"};
      const auto& mat = m_.material[mats[size_t(mat_of[e])]];
      m_.young[e] = mat.first;
      nus.insert(mat.second);
    }
    if (nus.size() != 1)
      throw InpError{UFC_STATUS_UNSUPPORTED, "several Poisson ratios (one nu per model)"};
    m_.nu = *nus.begin();
    for (auto& [eq, v] : m_.bc) {
      m_.bc_eq.push_back(eq);
      m_.bc_val.push_back(v);
    }
    for (auto& [eq, v] : m_.cload) {
      m_.cload_eq.push_back(eq);
      m_.cload_val.push_back(v);
    }
    for (auto& [lab, face, p] : m_.faces) {
      auto e = m_.elem_of.find(lab);
      if (e == m_.elem_of.end())
        throw InpError{UFC_STATUS_INVALID, "load on undefined element " + std::to_string(lab)};
      m_.face_elem.push_back(e->second);
      m_.face_id.push_back(face);
      m_.face_p.push_back(p);
    }
  }
};

void set(ufc_status* st, int code, const std::string& msg, int line = 0) {
  if (!st) return;
  std::memset(st, 0, sizeof *st);
  st->status_code = code;
  st->category = code;
  st->severity = code == UFC_STATUS_OK ? 0 : 3;
  st->line_number = line;
  std::snprintf(st->message, sizeof st->message, "%s", msg.c_str());
  std::snprintf(st->source, sizeof st->source, "%s", "ufc_ap_read_inp");
}

}  // namespace

extern "C" {

int ufc_ap_read_inp(const char* path, ufc_ap_model* model, ufc_status* st) {
  if (!path || !model) {
    set(st, UFC_STATUS_INVALID, "null argument");
    return UFC_STATUS_INVALID;
  }
  std::memset(model, 0, sizeof *model);
  try {
    auto m = std::make_unique<Model>();
    Reader(*m).run(path);
    model->n_nodes = int32_t(m->node_label.size());
    model->n_elems = int32_t(m->elem_label.size());
    model->node_label = m->node_label.data();
    model->elem_label = m->elem_label.data();
    model->coords = m->coords.data();
    model->conn = m->conn.data();
    model->young = m->young.data();
    model->nu = m->nu;
    model->variant = m->variant;
    model->n_bc = int32_t(m->bc_eq.size());
    model->bc_eq = m->bc_eq.data();
    model->bc_val = m->bc_val.data();
    model->n_cload = int32_t(m->cload_eq.size());
    model->cload_eq = m->cload_eq.data();
    model->cload_val = m->cload_val.data();
    model->n_face = int32_t(m->face_elem.size());
    model->face_elem = m->face_elem.data();
    model->face_id = m->face_id.data();
    model->face_p = m->face_p.data();
    model->impl = m.release();
    set(st, UFC_STATUS_OK, "read " + std::to_string(model->n_nodes) + " nodes, " +
                               std::to_string(model->n_elems) + " elements");
    return UFC_STATUS_OK;
  } catch (const InpError& e) {
    set(st, e.code, e.msg, e.line);
    return e.code;
  } catch (const std::bad_alloc&) {
    set(st, UFC_STATUS_MEM_ERROR, "host allocation failed");
    return UFC_STATUS_MEM_ERROR;
  } catch (const std::exception& e) {
    set(st, UFC_STATUS_INVALID, e.what());
    return UFC_STATUS_INVALID;
  }
}

int ufc_ap_free(ufc_ap_model* model, ufc_status* st) {
  if (model) {
    delete static_cast<Model*>(model->impl);
    std::memset(model, 0, sizeof *model);
  }
  set(st, UFC_STATUS_OK, "");
  return UFC_STATUS_OK;
}

}  // extern "C"
