#include "yaml_lite.hpp"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace shaw {
namespace yaml {

namespace {

const Node kNull;

std::string trim(const std::string& s) {
  size_t b = s.find_first_not_of(" \t\r");
  if (b == std::string::npos) return "";
  size_t e = s.find_last_not_of(" \t\r");
  return s.substr(b, e - b + 1);
}

std::string unquote(const std::string& s) {
  if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
    return s.substr(1, s.size() - 2);
  return s;
}

// strip a trailing comment: '#' at line start or after whitespace, outside quotes
std::string stripComment(const std::string& s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (q) { if (c == q) q = 0; continue; }
    if (c == '"' || c == '\'') { q = c; continue; }
    if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) return s.substr(0, i);
  }
  return s;
}

int bracketBalance(const std::string& s) {
  int d = 0;
  char q = 0;
  for (char c : s) {
    if (q) { if (c == q) q = 0; continue; }
    if (c == '"' || c == '\'') q = c;
    else if (c == '[' || c == '{') ++d;
    else if (c == ']' || c == '}') --d;
  }
  return d;
}

struct Line {
  int indent;
  std::string text;
  int lineno;
};

// position of the "key: value" separator (':' followed by space or end), or npos
size_t keySep(const std::string& s) {
  char q = 0;
  int d = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (q) { if (c == q) q = 0; continue; }
    if (c == '"' || c == '\'') { q = c; continue; }
    if (c == '[' || c == '{') ++d;
    if (c == ']' || c == '}') --d;
    if (c == ':' && d == 0 && (i + 1 == s.size() || s[i + 1] == ' ' || s[i + 1] == '\t')) return i;
  }
  return std::string::npos;
}

class FlowParser {
 public:
  FlowParser(const std::string& s, int lineno) : s_(s), line_(lineno) {}
  Node parseValue() {
    ws();
    if (i_ >= s_.size()) return Node();
    if (s_[i_] == '[') return parseSeq();
    if (s_[i_] == '{') return parseMap();
    return Node::scalar(parseScalar());
  }
  void expectEnd() {
    ws();
    if (i_ != s_.size()) fail("trailing characters in flow value");
  }

 private:
  void ws() { while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t')) ++i_; }
  [[noreturn]] void fail(const std::string& m) {
    throw std::runtime_error("yaml: line " + std::to_string(line_) + ": " + m);
  }
  std::string parseScalar() {
    ws();
    if (i_ < s_.size() && (s_[i_] == '"' || s_[i_] == '\'')) {
      const char q = s_[i_++];
      size_t b = i_;
      while (i_ < s_.size() && s_[i_] != q) ++i_;
      if (i_ >= s_.size()) fail("unterminated quote");
      std::string out = s_.substr(b, i_ - b);
      ++i_;
      return out;
    }
    size_t b = i_;
    while (i_ < s_.size() && s_[i_] != ',' && s_[i_] != ']' && s_[i_] != '}') {
      if (s_[i_] == ':' && (i_ + 1 == s_.size() || s_[i_ + 1] == ' ')) break;
      ++i_;
    }
    return trim(s_.substr(b, i_ - b));
  }
  Node parseSeq() {
    Node n = Node::sequence();
    ++i_;
    ws();
    if (i_ < s_.size() && s_[i_] == ']') { ++i_; return n; }
    while (true) {
      n.push(parseValue());
      ws();
      if (i_ >= s_.size()) fail("unterminated '['");
      if (s_[i_] == ',') { ++i_; continue; }
      if (s_[i_] == ']') { ++i_; return n; }
      fail("expected ',' or ']'");
    }
  }
  Node parseMap() {
    Node n = Node::map();
    ++i_;
    ws();
    if (i_ < s_.size() && s_[i_] == '}') { ++i_; return n; }
    while (true) {
      std::string key = parseScalar();
      ws();
      if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':' in flow map");
      ++i_;
      Node v = parseValue();
      n.set(key, std::move(v));
      ws();
      if (i_ >= s_.size()) fail("unterminated '{'");
      if (s_[i_] == ',') { ++i_; continue; }
      if (s_[i_] == '}') { ++i_; return n; }
      fail("expected ',' or '}'");
    }
  }
  const std::string& s_;
  size_t i_ = 0;
  int line_;
};

Node parseInline(const std::string& text, int lineno) {
  const std::string t = trim(text);
  if (t.empty()) return Node();
  if (t[0] == '[' || t[0] == '{') {
    FlowParser fp(t, lineno);
    Node n = fp.parseValue();
    fp.expectEnd();
    return n;
  }
  return Node::scalar(unquote(t));
}

bool isSeqItem(const std::string& t) { return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' '); }

class BlockParser {
 public:
  explicit BlockParser(std::vector<Line> lines) : L_(std::move(lines)) {}
  Node parseDocument() {
    if (L_.empty()) return Node::map();
    Node n = parseBlock(L_[0].indent);
    if (i_ != L_.size())
      throw std::runtime_error("yaml: line " + std::to_string(L_[i_].lineno) + ": bad indentation");
    return n;
  }

 private:
  Node parseBlock(int indent) {
    if (isSeqItem(L_[i_].text)) return parseSeq(indent);
    return parseMap(indent);
  }
  Node parseSeq(int indent) {
    Node n = Node::sequence();
    while (i_ < L_.size() && L_[i_].indent == indent && isSeqItem(L_[i_].text)) {
      const Line& ln = L_[i_];
      std::string rest = trim(ln.text.substr(1));
      ++i_;
      if (rest.empty()) {
        if (i_ < L_.size() && L_[i_].indent > indent) n.push(parseBlock(L_[i_].indent));
        else n.push(Node());
      } else if (rest[0] != '[' && rest[0] != '{' && keySep(rest) != std::string::npos) {
        // "- key: value" : a mapping whose first entry sits on the item line
        const int sub = indent + 1 + (int)(ln.text.size() - 1 - trim(ln.text.substr(1)).size());
        --i_;
        L_[i_].indent = sub;
        L_[i_].text = rest;
        n.push(parseMap(sub));
      } else {
        n.push(parseInline(rest, ln.lineno));
      }
    }
    return n;
  }
  Node parseMap(int indent) {
    Node n = Node::map();
    while (i_ < L_.size() && L_[i_].indent == indent && !isSeqItem(L_[i_].text)) {
      const Line ln = L_[i_];
      const size_t sep = keySep(ln.text);
      if (sep == std::string::npos)
        throw std::runtime_error("yaml: line " + std::to_string(ln.lineno) + ": expected 'key: value'");
      const std::string key = unquote(trim(ln.text.substr(0, sep)));
      const std::string rest = trim(ln.text.substr(sep + 1));
      ++i_;
      if (!rest.empty()) {
        n.set(key, parseInline(rest, ln.lineno));
      } else if (i_ < L_.size() && L_[i_].indent > indent) {
        n.set(key, parseBlock(L_[i_].indent));
      } else if (i_ < L_.size() && L_[i_].indent == indent && isSeqItem(L_[i_].text)) {
        n.set(key, parseSeq(indent));
      } else {
        n.set(key, Node());
      }
    }
    if (i_ < L_.size() && L_[i_].indent > indent)
      throw std::runtime_error("yaml: line " + std::to_string(L_[i_].lineno) + ": unexpected indentation");
    return n;
  }
  std::vector<Line> L_;
  size_t i_ = 0;
};

}  // namespace

const Node& Node::operator[](const std::string& key) const {
  if (kind_ != Kind::Map) return kNull;
  for (const auto& e : map_)
    if (e.first == key) return e.second;
  return kNull;
}

void Node::set(std::string k, Node v) {
  for (auto& e : map_)
    if (e.first == k) { e.second = std::move(v); return; }  // later duplicate wins
  map_.emplace_back(std::move(k), std::move(v));
}

const std::string& Node::str() const {
  if (kind_ != Kind::Scalar) throw std::runtime_error("yaml: node is not a scalar");
  return scalar_;
}

double Node::asDouble() const {
  const std::string& s = str();
  errno = 0;
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || trim(end).size() != 0) throw std::runtime_error("yaml: bad number '" + s + "'");
  return v;
}

long long Node::asInt() const {
  const std::string& s = str();
  char* end = nullptr;
  const long long v = std::strtoll(s.c_str(), &end, 10);
  if (end == s.c_str() || trim(end).size() != 0) throw std::runtime_error("yaml: bad integer '" + s + "'");
  return v;
}

size_t Node::asSize() const {
  const long long v = asInt();
  if (v < 0) throw std::runtime_error("yaml: negative value for an unsigned entry '" + str() + "'");
  return (size_t)v;
}

bool Node::asBool() const {
  const std::string& s = str();
  static const char* T[] = {"true", "True", "TRUE", "yes", "Yes", "YES", "on", "On", "ON", "y", "Y"};
  static const char* F[] = {"false", "False", "FALSE", "no", "No", "NO", "off", "Off", "OFF", "n", "N"};
  for (auto t : T) if (s == t) return true;
  for (auto f : F) if (s == f) return false;
  throw std::runtime_error("yaml: bad bool '" + s + "'");
}

std::vector<double> Node::asDoubleVector() const {
  if (kind_ != Kind::Sequence) throw std::runtime_error("yaml: node is not a sequence");
  std::vector<double> out;
  for (const auto& n : seq_) out.push_back(n.asDouble());
  return out;
}

Node parse(const std::string& text) {
  std::vector<Line> lines;
  std::istringstream in(text);
  std::string raw;
  int lineno = 0;
  std::string pending;
  int pendingIndent = 0, pendingLine = 0;
  while (std::getline(in, raw)) {
    ++lineno;
    for (char& c : raw) if (c == '\t') c = ' ';
    std::string s = stripComment(raw);
    if (trim(s).empty()) continue;
    if (!pending.empty()) {
      pending += " " + trim(s);
      if (bracketBalance(pending) <= 0) { lines.push_back({pendingIndent, pending, pendingLine}); pending.clear(); }
      continue;
    }
    if (trim(s) == "---") continue;
    const int indent = (int)s.find_first_not_of(' ');
    std::string t = trim(s);
    if (bracketBalance(t) > 0) { pending = t; pendingIndent = indent; pendingLine = lineno; continue; }
    lines.push_back({indent, t, lineno});
  }
  if (!pending.empty()) throw std::runtime_error("yaml: unterminated flow collection");
  BlockParser bp(std::move(lines));
  return bp.parseDocument();
}

Node loadFile(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("bad file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

}  // namespace yaml
}  // namespace shaw
