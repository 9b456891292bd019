#include "xml_lite.hpp"

#include <cctype>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace hpo {

const XNode* XNode::child(const std::string& n) const {
  for (auto& c : children)
    if (c->name == n) return c.get();
  return nullptr;
}

std::vector<const XNode*> XNode::all(const std::string& n) const {
  std::vector<const XNode*> out;
  for (auto& c : children)
    if (c->name == n) out.push_back(c.get());
  return out;
}

std::string XNode::get(const std::string& a, const std::string& dflt) const {
  auto it = attr.find(a);
  return it == attr.end() ? dflt : it->second;
}

namespace {

class Parser {
 public:
  explicit Parser(const std::string& s) : s_(s) {}

  std::unique_ptr<XNode> document() {
    skip_misc();
    if (!at('<')) fail("expected the root element");
    auto root = element();
    skip_misc();
    if (p_ != s_.size()) fail("content after the root element");
    return root;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const std::string& what) const {
    size_t line = 1;
    for (size_t i = 0; i < p_ && i < s_.size(); ++i) line += (s_[i] == '\n');
    throw std::runtime_error("XML line " + std::to_string(line) + ": " + what);
  }
  bool at(char c) const { return p_ < s_.size() && s_[p_] == c; }
  bool starts(const char* t) const { return s_.compare(p_, std::char_traits<char>::length(t), t) == 0; }
  void ws() {
    while (p_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[p_]))) ++p_;
  }
  void skip_until(const char* t) {
    size_t q = s_.find(t, p_);
    if (q == std::string::npos) fail(std::string("unterminated construct, expected ") + t);
    p_ = q + std::char_traits<char>::length(t);
  }
  // whitespace, comments, declarations and processing instructions between elements
  void skip_misc() {
    for (;;) {
      ws();
      if (starts("<!--")) skip_until("-->");
      else if (starts("<?")) skip_until("?>");
      else if (starts("<!")) skip_until(">");
      else return;
    }
  }
  std::string name() {
    size_t b = p_;
    while (p_ < s_.size() && (std::isalnum(static_cast<unsigned char>(s_[p_])) || s_[p_] == '_' || s_[p_] == '-' ||
                              s_[p_] == ':' || s_[p_] == '.'))
      ++p_;
    if (b == p_) fail("expected a name");
    return s_.substr(b, p_ - b);
  }
  static std::string decode(const std::string& t) {
    std::string o;
    o.reserve(t.size());
    for (size_t i = 0; i < t.size(); ++i) {
      if (t[i] != '&') {
        o += t[i];
        continue;
      }
      size_t e = t.find(';', i);
      std::string ent = e == std::string::npos ? "" : t.substr(i + 1, e - i - 1);
      if (ent == "lt") o += '<';
      else if (ent == "gt") o += '>';
      else if (ent == "amp") o += '&';
      else if (ent == "quot") o += '"';
      else if (ent == "apos") o += '\'';
      else {
        o += '&';
        continue;
      }
      i = e;
    }
    return o;
  }
  std::unique_ptr<XNode> element() {
    ++p_;  // '<'
    auto n = std::make_unique<XNode>();
    n->name = name();
    for (;;) {
      ws();
      if (starts("/>")) {
        p_ += 2;
        return n;
      }
      if (at('>')) {
        ++p_;
        break;
      }
      std::string a = name();
      ws();
      if (!at('=')) fail("expected '=' after attribute " + a);
      ++p_;
      ws();
      if (!at('"') && !at('\'')) fail("expected a quoted value for attribute " + a);
      char q = s_[p_++];
      size_t e = s_.find(q, p_);
      if (e == std::string::npos) fail("unterminated attribute value");
      n->attr[a] = decode(s_.substr(p_, e - p_));
      p_ = e + 1;
    }
    for (;;) {
      size_t lt = s_.find('<', p_);
      if (lt == std::string::npos) fail("unterminated element " + n->name);
      n->text += decode(s_.substr(p_, lt - p_));
      p_ = lt;
      if (starts("<!--")) {
        skip_until("-->");
      } else if (starts("<?")) {
        skip_until("?>");
      } else if (starts("</")) {
        p_ += 2;
        std::string c = name();
        if (c != n->name) fail("mismatched closing tag </" + c + "> for <" + n->name + ">");
        ws();
        if (!at('>')) fail("expected '>'");
        ++p_;
        return n;
      } else {
        n->children.push_back(element());
      }
    }
  }
};

}  // namespace

std::unique_ptr<XNode> xml_parse(const std::string& text) { return Parser(text).document(); }

std::unique_ptr<XNode> xml_load_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return xml_parse(ss.str());
}

}  // namespace hpo
