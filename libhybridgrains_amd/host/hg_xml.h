// hg_xml.h -- a small XML reader for the scene files: elements, attributes,
// comments and declarations; text content is ignored (the scene schemas keep
// everything in attributes).  Element and attribute order is preserved, as the
// scene parsers iterate siblings in file order.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace hg {
namespace xml {

struct Node {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::unique_ptr<Node>> children;

  // First attribute with this name, or nullptr.
  const std::string *attr(const std::string &key) const;
  // First child element with this name, or nullptr.
  const Node *child(const std::string &key) const;
  // All child elements with this name, in file order.
  std::vector<const Node *> children_named(const std::string &key) const;
};

// Parses a whole document; returns the synthetic document node whose children
// are the top-level elements, or nullptr with a message in *error.
std::unique_ptr<Node> parse(const std::string &text, std::string *error);
std::unique_ptr<Node> parse_file(const std::string &path, std::string *error);

}  // namespace xml
}  // namespace hg
