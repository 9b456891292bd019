// Minimal XML reader for Cantera CTML mechanism files: elements, attributes and text.
// Comments, the XML declaration and processing instructions are skipped; entity references
// &lt; &gt; &amp; &quot; &apos; are decoded. No DTDs, namespaces or CDATA.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace hpo {

struct XNode {
  std::string name;
  std::map<std::string, std::string> attr;
  std::string text;   // concatenated character data directly inside this element
  std::vector<std::unique_ptr<XNode>> children;

  const XNode* child(const std::string& n) const;            // first child named n, or null
  std::vector<const XNode*> all(const std::string& n) const;  // every child named n
  std::string get(const std::string& a, const std::string& dflt = "") const;
  bool has(const std::string& a) const { return attr.count(a) != 0; }
};

// Parses a document; returns its root element. Throws std::runtime_error with the line number
// on malformed input.
std::unique_ptr<XNode> xml_parse(const std::string& text);
std::unique_ptr<XNode> xml_load_file(const std::string& path);

}  // namespace hpo
