// A small YAML reader for NCAR-format mechanism and scenario files.
//
// Scope: block mappings, block sequences (including sequences indented at the
// same column as their parent key), flow sequences/mappings on one logical
// line, plain / single-quoted / double-quoted scalars, multi-line plain
// scalars and comments.  Every scalar is kept as its text (no type guessing),
// so a species named NO stays "NO" instead of becoming a boolean.  Mapping
// keys keep document order, which is the order yaml-cpp iterates them in the
// reference loader (src/core/TChem_KineticModelData.cpp:110-131,190-253).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace tchem_amd {
namespace yaml {

struct Node {
  enum Kind { Null, Scalar, Sequence, Map };
  Kind kind = Null;
  std::string text;                                   // Scalar
  std::vector<Node> items;                            // Sequence
  std::vector<std::pair<std::string, Node>> entries;  // Map (document order)

  bool isNull() const { return kind == Null; }
  bool isMap() const { return kind == Map; }
  bool isSeq() const { return kind == Sequence; }
  bool isScalar() const { return kind == Scalar; }

  // Map lookup; returns nullptr when absent or when this node is not a map.
  const Node* find(const std::string& key) const;
  bool has(const std::string& key) const { return find(key) != nullptr; }
  const Node& operator[](const std::string& key) const;  // throws if absent

  // Number of children (sequence items or map entries); 0 for scalars/null.
  size_t size() const {
    return kind == Sequence ? items.size() : (kind == Map ? entries.size() : 0);
  }

  const std::string& str() const;  // scalar text; throws otherwise
  double asDouble() const;         // strtod on the scalar text; throws on junk
  int asInt() const;
};

struct ParseError : public std::runtime_error {
  using std::runtime_error::runtime_error;
};

Node parse(const std::string& text);
Node parseFile(const std::string& path);

}  // namespace yaml
}  // namespace tchem_amd
