// A small YAML-subset reader for SHAW input files.
//
// Supports what the reference's inputs use (src/shared/parser/*.hpp read them
// through yaml-cpp 0.7.0): block mappings by indentation, block sequences
// ("- item", also at the parent key's indentation), flow mappings
// "{a: 1, b: [1, 2]}" and flow sequences "[...]" (possibly spanning lines),
// single/double quoted scalars, "#" comments, and scalar conversion to
// bool/int/size_t/double/string with yaml-cpp's accepted spellings.
#pragma once

#include <cstddef>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace shaw {
namespace yaml {

class Node {
 public:
  enum class Kind { Null, Scalar, Sequence, Map };

  Node() = default;
  static Node scalar(std::string v) { Node n; n.kind_ = Kind::Scalar; n.scalar_ = std::move(v); return n; }
  static Node sequence() { Node n; n.kind_ = Kind::Sequence; return n; }
  static Node map() { Node n; n.kind_ = Kind::Map; return n; }

  Kind kind() const { return kind_; }
  bool isNull() const { return kind_ == Kind::Null; }
  bool isScalar() const { return kind_ == Kind::Scalar; }
  bool isSequence() const { return kind_ == Kind::Sequence; }
  bool isMap() const { return kind_ == Kind::Map; }
  explicit operator bool() const { return kind_ != Kind::Null; }

  // map access; a missing key gives a Null node (as yaml-cpp's operator[])
  const Node& operator[](const std::string& key) const;
  const std::vector<Node>& items() const { return seq_; }
  const std::vector<std::pair<std::string, Node>>& entries() const { return map_; }
  size_t size() const { return kind_ == Kind::Sequence ? seq_.size() : map_.size(); }

  const std::string& str() const;
  double asDouble() const;
  long long asInt() const;
  size_t asSize() const;
  bool asBool() const;
  std::vector<double> asDoubleVector() const;

  void push(Node n) { seq_.push_back(std::move(n)); }
  void set(std::string k, Node v);

 private:
  Kind kind_ = Kind::Null;
  std::string scalar_;
  std::vector<Node> seq_;
  std::vector<std::pair<std::string, Node>> map_;
};

// Parse a document from text / from a file. Throws std::runtime_error.
Node parse(const std::string& text);
Node loadFile(const std::string& path);

}  // namespace yaml
}  // namespace shaw
