// Minimal YAML reader for Cantera mechanism files and the integrator's config.yaml.
//
// yaml-cpp is what the reference links (src/Integrators/CMakeLists.txt:20); it is not
// available here, so this covers the subset those files use: block mappings and
// sequences by indentation, flow sequences [..] and flow mappings {..} (also across
// lines), plain and quoted scalars, comments, and literal/folded block scalars.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace hpo {

struct YNode {
  enum Kind { Null, Scalar, Seq, Map } kind = Null;
  std::string scalar;
  std::vector<YNode> seq;
  std::vector<std::pair<std::string, YNode>> map;  // insertion order preserved

  bool is_null() const { return kind == Null; }
  bool has(const std::string& k) const;
  const YNode& operator[](const std::string& k) const;  // throws if missing
  const YNode* find(const std::string& k) const;
  const YNode& operator[](size_t i) const { return seq.at(i); }
  size_t size() const { return kind == Seq ? seq.size() : (kind == Map ? map.size() : 0); }
  double as_double() const;
  long as_long() const;
  const std::string& as_string() const;
};

YNode yaml_parse(const std::string& text);
YNode yaml_load_file(const std::string& path);

struct YamlError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

}  // namespace hpo
