#include "cti_lite.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace hpo {

const CNode* CNode::kw(const std::string& k) const {
  for (auto& kv : kwargs)
    if (kv.first == k) return kv.second.get();
  return nullptr;
}

double CNode::as_double() const {
  if (kind != Num) throw std::runtime_error("CTI: expected a number" + (kind == Str ? ", got '" + str + "'" : std::string()));
  return num;
}

const std::string& CNode::as_string() const {
  if (kind != Str) throw std::runtime_error("CTI: expected a string");
  return str;
}

namespace {

struct Tok {
  enum Kind { End, Num, Str, Name, Punct } kind = End;
  std::string text;
  double num = 0.0;
  int line = 0;
};

class Lexer {
 public:
  explicit Lexer(const std::string& s) : s_(s) {}

  Tok next() {
    skip();
    Tok t;
    t.line = line_;
    if (p_ >= s_.size()) return t;
    const char c = s_[p_];
    if (c == '\'' || c == '"') {
      t.kind = Tok::Str;
      t.text = string_lit();
      // adjacent string literals are one string, as in Python
      for (;;) {
        const size_t save = p_;
        const int save_line = line_;
        skip();
        if (p_ < s_.size() && (s_[p_] == '\'' || s_[p_] == '"')) {
          t.text += string_lit();
        } else {
          p_ = save;
          line_ = save_line;
          break;
        }
      }
      return t;
    }
    if (std::isdigit(static_cast<unsigned char>(c)) || ((c == '-' || c == '+' || c == '.') && number_follows())) {
      const char* b = s_.c_str() + p_;
      char* e = nullptr;
      t.num = std::strtod(b, &e);
      if (e == b) error("bad number");
      p_ += static_cast<size_t>(e - b);
      t.kind = Tok::Num;
      return t;
    }
    if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
      size_t b = p_;
      while (p_ < s_.size() && (std::isalnum(static_cast<unsigned char>(s_[p_])) || s_[p_] == '_' || s_[p_] == '.')) ++p_;
      t.kind = Tok::Name;
      t.text = s_.substr(b, p_ - b);
      return t;
    }
    if (c == '(' || c == ')' || c == '[' || c == ']' || c == ',' || c == '=') {
      ++p_;
      t.kind = Tok::Punct;
      t.text = std::string(1, c);
      return t;
    }
    error(std::string("unsupported character '") + c + "'");
    return t;
  }

  [[noreturn]] void error(const std::string& what) const {
    throw std::runtime_error("CTI line " + std::to_string(line_) + ": " + what);
  }

 private:
  void skip() {
    while (p_ < s_.size()) {
      const char c = s_[p_];
      if (c == '\n') {
        ++line_;
        ++p_;
      } else if (c == '#') {
        while (p_ < s_.size() && s_[p_] != '\n') ++p_;
      } else if (c == '\\' && p_ + 1 < s_.size() && s_[p_ + 1] == '\n') {
        p_ += 2;
        ++line_;
      } else if (std::isspace(static_cast<unsigned char>(c))) {
        ++p_;
      } else {
        break;
      }
    }
  }
  bool number_follows() const {
    size_t q = p_;
    if (s_[q] == '-' || s_[q] == '+') ++q;
    if (q < s_.size() && s_[q] == '.') ++q;
    return q < s_.size() && std::isdigit(static_cast<unsigned char>(s_[q]));
  }
  std::string string_lit() {
    const char q = s_[p_];
    const bool triple = p_ + 2 < s_.size() && s_[p_ + 1] == q && s_[p_ + 2] == q;
    p_ += triple ? 3 : 1;
    std::string out;
    for (;;) {
      if (p_ >= s_.size()) error("unterminated string");
      const char c = s_[p_];
      if (triple) {
        if (c == q && p_ + 2 < s_.size() && s_[p_ + 1] == q && s_[p_ + 2] == q) {
          p_ += 3;
          return out;
        }
      } else if (c == q) {
        ++p_;
        return out;
      } else if (c == '\n') {
        error("newline in string");
      }
      if (c == '\\' && p_ + 1 < s_.size()) {
        out += s_[p_ + 1];
        p_ += 2;
        continue;
      }
      if (c == '\n') ++line_;
      out += c;
      ++p_;
    }
  }

  const std::string& s_;
  size_t p_ = 0;
  int line_ = 1;
};

class Parser {
 public:
  explicit Parser(const std::string& s) : lx_(s) { adv(); }

  std::vector<std::unique_ptr<CNode>> file() {
    std::vector<std::unique_ptr<CNode>> out;
    while (t_.kind != Tok::End) {
      if (t_.kind == Tok::Str) {   // a string statement (the file's docstring) is skipped
        adv();
        continue;
      }
      if (t_.kind != Tok::Name) lx_.error("expected a statement");
      std::string name = t_.text;
      adv();
      if (is('=')) {   // a top-level assignment: its value is read and dropped
        adv();
        (void)value();
        continue;
      }
      if (!is('(')) lx_.error("expected '(' after " + name);
      out.push_back(call(name));
    }
    return out;
  }

 private:
  void adv() { t_ = lx_.next(); }
  bool is(char c) const { return t_.kind == Tok::Punct && t_.text[0] == c; }
  void expect(char c) {
    if (!is(c)) lx_.error(std::string("expected '") + c + "'");
    adv();
  }

  std::unique_ptr<CNode> call(const std::string& name) {
    auto n = std::make_unique<CNode>();
    n->kind = CNode::Call;
    n->str = name;
    expect('(');
    while (!is(')')) {
      if (t_.kind == Tok::Name) {
        std::string nm = t_.text;
        adv();
        if (is('=')) {
          adv();
          n->kwargs.emplace_back(nm, value());
        } else {
          n->items.push_back(after_name(nm));
        }
      } else {
        n->items.push_back(value());
      }
      if (is(',')) adv();
      else if (!is(')')) lx_.error("expected ',' or ')' in " + name + "(...)");
    }
    adv();
    return n;
  }

  std::unique_ptr<CNode> after_name(const std::string& nm) {
    if (is('(')) return call(nm);
    auto n = std::make_unique<CNode>();
    n->kind = CNode::Name;
    n->str = nm;
    return n;
  }

  std::unique_ptr<CNode> seq(CNode::Kind kind, char close) {
    auto n = std::make_unique<CNode>();
    n->kind = kind;
    adv();
    while (!is(close)) {
      n->items.push_back(value());
      if (is(',')) adv();
      else if (!is(close)) lx_.error(std::string("expected ',' or '") + close + "'");
    }
    adv();
    return n;
  }

  std::unique_ptr<CNode> value() {
    if (t_.kind == Tok::Num) {
      auto n = std::make_unique<CNode>();
      n->kind = CNode::Num;
      n->num = t_.num;
      adv();
      return n;
    }
    if (t_.kind == Tok::Str) {
      auto n = std::make_unique<CNode>();
      n->kind = CNode::Str;
      n->str = t_.text;
      adv();
      return n;
    }
    if (t_.kind == Tok::Name) {
      std::string nm = t_.text;
      adv();
      return after_name(nm);
    }
    if (is('[')) return seq(CNode::List, ']');
    if (is('(')) {
      // a parenthesised value or a tuple
      auto n = seq(CNode::Tuple, ')');
      return n;
    }
    lx_.error("expected a value");
  }

  Lexer lx_;
  Tok t_;
};

}  // namespace

std::vector<std::unique_ptr<CNode>> cti_parse(const std::string& text) {
  Parser p(text);
  return p.file();
}

std::vector<std::unique_ptr<CNode>> cti_load_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return cti_parse(ss.str());
}

}  // namespace hpo
