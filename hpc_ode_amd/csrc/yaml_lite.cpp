#include "yaml_lite.hpp"

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace hpo {

bool YNode::has(const std::string& k) const { return find(k) != nullptr; }

const YNode* YNode::find(const std::string& k) const {
  if (kind != Map) return nullptr;
  for (const auto& kv : map)
    if (kv.first == k) return &kv.second;
  return nullptr;
}

const YNode& YNode::operator[](const std::string& k) const {
  const YNode* n = find(k);
  if (!n) throw YamlError("missing key '" + k + "'");
  return *n;
}

double YNode::as_double() const {
  if (kind != Scalar) throw YamlError("expected a scalar number");
  const char* s = scalar.c_str();
  char* end = nullptr;
  double v = std::strtod(s, &end);
  if (end == s) throw YamlError("not a number: '" + scalar + "'");
  return v;
}

long YNode::as_long() const { return static_cast<long>(as_double()); }

const std::string& YNode::as_string() const {
  if (kind != Scalar) throw YamlError("expected a scalar string");
  return scalar;
}

namespace {

struct Line {
  int indent;
  std::string text;  // without indentation, comments stripped, flow continuations joined
  int lineno;
};

std::string rstrip(const std::string& s) {
  size_t e = s.find_last_not_of(" \t\r");
  return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}

std::string strip(const std::string& s) {
  size_t b = s.find_first_not_of(" \t\r");
  if (b == std::string::npos) return std::string();
  size_t e = s.find_last_not_of(" \t\r");
  return s.substr(b, e - b + 1);
}

// Removes a trailing comment (a '#' at line start or preceded by whitespace, outside quotes).
std::string strip_comment(const std::string& s) {
  bool sq = false, dq = false;
  for (size_t i = 0; i < s.size(); ++i) {
    char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (c == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t'))
      return s.substr(0, i);
  }
  return s;
}

int flow_depth_delta(const std::string& s) {
  int d = 0;
  bool sq = false, dq = false;
  for (char c : s) {
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (!sq && !dq) {
      if (c == '[' || c == '{') ++d;
      else if (c == ']' || c == '}') --d;
    }
  }
  return d;
}

std::vector<Line> logical_lines(const std::string& text) {
  std::vector<std::string> raw;
  {
    std::istringstream is(text);
    std::string l;
    while (std::getline(is, l)) raw.push_back(l);
  }
  std::vector<Line> out;
  for (size_t i = 0; i < raw.size(); ++i) {
    std::string l = rstrip(raw[i]);
    std::string nc = rstrip(strip_comment(l));
    if (strip(nc).empty()) continue;
    if (strip(nc) == "---" || strip(nc) == "...") continue;
    int ind = static_cast<int>(nc.find_first_not_of(' '));
    std::string body = nc.substr(ind);
    // block scalar: key: | or key: >  (also "- |")
    std::string tb = strip(body);
    bool block = false;
    {
      size_t p = tb.rfind(':');
      std::string tail = (p != std::string::npos) ? strip(tb.substr(p + 1)) : std::string();
      if (tb.size() >= 1 && (tail == "|" || tail == "|-" || tail == "|+" || tail == ">" ||
                             tail == ">-" || tail == ">+"))
        block = true;
    }
    if (block) {
      // gather following lines indented deeper than this one
      std::string acc;
      size_t j = i + 1;
      int bind = -1;
      for (; j < raw.size(); ++j) {
        std::string r = rstrip(raw[j]);
        if (strip(r).empty()) {
          acc += "\n";
          continue;
        }
        int ri = static_cast<int>(r.find_first_not_of(' '));
        if (ri <= ind) break;
        if (bind < 0) bind = ri;
        acc += r.substr(std::min<int>(bind, ri)) + "\n";
      }
      size_t p = body.rfind(':');
      std::string key = body.substr(0, p);
      // encode as a quoted scalar so the block parser treats it as text
      std::string esc;
      for (char c : acc) esc += (c == '"' ? '\'' : (c == '\n' ? ' ' : c));
      out.push_back({ind, key + ": \"" + esc + "\"", static_cast<int>(i + 1)});
      i = j - 1;
      continue;
    }
    int depth = flow_depth_delta(body);
    size_t j = i;
    while (depth > 0 && j + 1 < raw.size()) {
      ++j;
      std::string cont = strip(strip_comment(rstrip(raw[j])));
      body += " " + cont;
      depth += flow_depth_delta(cont);
    }
    if (depth != 0) throw YamlError("unbalanced flow collection at line " + std::to_string(i + 1));
    out.push_back({ind, body, static_cast<int>(i + 1)});
    i = j;
  }
  return out;
}

YNode make_scalar(std::string s) {
  YNode n;
  s = strip(s);
  if (s.empty() || s == "~" || s == "null") {
    n.kind = YNode::Null;
    return n;
  }
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
    s = s.substr(1, s.size() - 2);
  n.kind = YNode::Scalar;
  n.scalar = s;
  return n;
}

struct Flow {
  const std::string& s;
  size_t p = 0;
  explicit Flow(const std::string& str) : s(str) {}
  void ws() {
    while (p < s.size() && (s[p] == ' ' || s[p] == '\t')) ++p;
  }
  YNode value() {
    ws();
    if (p >= s.size()) return YNode();
    if (s[p] == '[') return seq();
    if (s[p] == '{') return mapping();
    return scalar_until(",]}");
  }
  YNode scalar_until(const char* stops) {
    ws();
    size_t b = p;
    if (p < s.size() && (s[p] == '"' || s[p] == '\'')) {
      char q = s[p++];
      while (p < s.size() && s[p] != q) ++p;
      ++p;
      return make_scalar(s.substr(b, p - b));
    }
    while (p < s.size() && !std::strchr(stops, s[p])) ++p;
    return make_scalar(s.substr(b, p - b));
  }
  YNode seq() {
    YNode n;
    n.kind = YNode::Seq;
    ++p;  // [
    ws();
    if (p < s.size() && s[p] == ']') {
      ++p;
      return n;
    }
    for (;;) {
      n.seq.push_back(value());
      ws();
      if (p >= s.size()) throw YamlError("unterminated flow sequence");
      if (s[p] == ',') {
        ++p;
        continue;
      }
      if (s[p] == ']') {
        ++p;
        break;
      }
      throw YamlError("bad flow sequence near: " + s.substr(p));
    }
    return n;
  }
  YNode mapping() {
    YNode n;
    n.kind = YNode::Map;
    ++p;  // {
    ws();
    if (p < s.size() && s[p] == '}') {
      ++p;
      return n;
    }
    for (;;) {
      YNode k = scalar_until(":,}");
      ws();
      if (p >= s.size() || s[p] != ':') throw YamlError("bad flow mapping near: " + s.substr(p));
      ++p;
      YNode v = value();
      n.map.emplace_back(k.scalar, v);
      ws();
      if (p >= s.size()) throw YamlError("unterminated flow mapping");
      if (s[p] == ',') {
        ++p;
        continue;
      }
      if (s[p] == '}') {
        ++p;
        break;
      }
      throw YamlError("bad flow mapping near: " + s.substr(p));
    }
    return n;
  }
};

YNode parse_inline_value(const std::string& v) {
  std::string t = strip(v);
  if (!t.empty() && (t[0] == '[' || t[0] == '{')) {
    Flow f(t);
    return f.value();
  }
  return make_scalar(t);
}

// Finds "key: " separator outside quotes/flow; returns npos if none.
size_t find_key_sep(const std::string& s) {
  bool sq = false, dq = false;
  int depth = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (!sq && !dq) {
      if (c == '[' || c == '{') ++depth;
      else if (c == ']' || c == '}') --depth;
      else if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
    }
  }
  return std::string::npos;
}

struct Parser {
  std::vector<Line> L;
  size_t i = 0;

  YNode block(int indent) {
    if (i >= L.size() || L[i].indent < indent) return YNode();
    int ind = L[i].indent;
    const std::string& t = L[i].text;
    if (t == "-" || t.rfind("- ", 0) == 0) return sequence(ind);
    if (find_key_sep(t) != std::string::npos) return mapping(ind);
    // lone scalar / flow value
    YNode v = parse_inline_value(t);
    ++i;
    return v;
  }

  YNode sequence(int ind) {
    YNode n;
    n.kind = YNode::Seq;
    while (i < L.size() && L[i].indent == ind && (L[i].text == "-" || L[i].text.rfind("- ", 0) == 0)) {
      std::string rest = L[i].text == "-" ? std::string() : L[i].text.substr(2);
      size_t off = rest.find_first_not_of(' ');
      if (rest.empty() || off == std::string::npos) {
        ++i;
        n.seq.push_back(block(ind + 1));
        continue;
      }
      int cind = ind + 2 + static_cast<int>(off);
      rest = rest.substr(off);
      // rewrite the current line as the first line of the item at deeper indentation
      L[i].indent = cind;
      L[i].text = rest;
      n.seq.push_back(block(cind));
    }
    return n;
  }

  YNode mapping(int ind) {
    YNode n;
    n.kind = YNode::Map;
    while (i < L.size() && L[i].indent == ind) {
      const std::string t = L[i].text;
      if (t == "-" || t.rfind("- ", 0) == 0) break;
      size_t sep = find_key_sep(t);
      if (sep == std::string::npos)
        throw YamlError("expected 'key: value' at line " + std::to_string(L[i].lineno));
      std::string key = strip(t.substr(0, sep));
      if (key.size() >= 2 && (key.front() == '"' || key.front() == '\''))
        key = key.substr(1, key.size() - 2);
      std::string val = strip(t.substr(sep + 1));
      ++i;
      if (!val.empty()) {
        n.map.emplace_back(key, parse_inline_value(val));
      } else if (i < L.size() && (L[i].indent > ind ||
                                  (L[i].indent == ind && (L[i].text == "-" || L[i].text.rfind("- ", 0) == 0)))) {
        n.map.emplace_back(key, block(L[i].indent));
      } else {
        n.map.emplace_back(key, YNode());
      }
    }
    return n;
  }
};

}  // namespace

YNode yaml_parse(const std::string& text) {
  Parser p;
  p.L = logical_lines(text);
  if (p.L.empty()) return YNode();
  YNode root = p.block(p.L[0].indent);
  if (p.i != p.L.size())
    throw YamlError("unexpected content at line " + std::to_string(p.L[p.i].lineno));
  return root;
}

YNode yaml_load_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw YamlError("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return yaml_parse(ss.str());
}

}  // namespace hpo
