// Reader for Cantera's CTI input format (the .cti files of the reference's
// resources/mechanisms/): a file of Python call statements such as
//   units(length='cm', time='s', quantity='mol', act_energy='cal/mol')
//   species(name='H2', atoms='H:2', thermo=(NASA([200.0, 1000.0], [...]), NASA(...)))
//   falloff_reaction('2 OH (+ M) <=> H2O2 (+ M)', kf=[...], kf0=[...], falloff=Troe(A=0.7346, ...))
// Only the literal subset those files use is read: numbers, strings (single, double and
// triple quoted, adjacent ones joined), names, lists, tuples and calls with positional and
// keyword arguments. Comments (#) are skipped. Nothing is evaluated.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace hpo {

struct CNode {
  enum Kind { Num, Str, Name, List, Tuple, Call } kind = Num;
  double num = 0.0;
  std::string str;                                   // Str: the text; Name / Call: the name
  std::vector<std::unique_ptr<CNode>> items;         // List / Tuple items, Call positional args
  std::vector<std::pair<std::string, std::unique_ptr<CNode>>> kwargs;   // Call keyword args

  const CNode* kw(const std::string& k) const;       // keyword argument or nullptr
  const CNode* arg(size_t i) const { return i < items.size() ? items[i].get() : nullptr; }
  double as_double() const;                          // Num (throws otherwise)
  const std::string& as_string() const;              // Str (throws otherwise)
};

// The top-level statements of a CTI file, in order (each a Call).
std::vector<std::unique_ptr<CNode>> cti_parse(const std::string& text);
std::vector<std::unique_ptr<CNode>> cti_load_file(const std::string& path);

}  // namespace hpo
