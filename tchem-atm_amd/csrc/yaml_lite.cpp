#include "yaml_lite.hpp"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace tchem_amd {
namespace yaml {

namespace {

struct Line {
  int indent;
  std::string text;  // content without indentation, comment or trailing blanks
  int lineno;
};

std::string rtrim(const std::string& s) {
  size_t e = s.find_last_not_of(" \t\r\n");
  return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}

std::string trim(const std::string& s) {
  size_t b = s.find_first_not_of(" \t\r\n");
  if (b == std::string::npos) return std::string();
  size_t e = s.find_last_not_of(" \t\r\n");
  return s.substr(b, e - b + 1);
}

// Remove a trailing comment: '#' at the start or after whitespace, outside
// quotes.
std::string stripComment(const std::string& s) {
  bool sq = false, dq = false;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (c == '#' && !sq && !dq && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t'))
      return s.substr(0, i);
  }
  return s;
}

std::vector<Line> splitLines(const std::string& text) {
  std::vector<Line> out;
  std::istringstream is(text);
  std::string raw;
  int lineno = 0;
  while (std::getline(is, raw)) {
    ++lineno;
    std::string s = rtrim(stripComment(raw));
    size_t b = s.find_first_not_of(' ');
    if (b == std::string::npos) continue;
    if (s[b] == '\t') throw ParseError("tab indentation at line " + std::to_string(lineno));
    std::string content = s.substr(b);
    if (b == 0 && (content == "---" || content == "...")) continue;
    out.push_back(Line{int(b), content, lineno});
  }
  return out;
}

std::string unquote(const std::string& s, int lineno) {
  if (s.size() >= 2 && s.front() == '\'' && s.back() == '\'') {
    std::string r;
    for (size_t i = 1; i + 1 < s.size(); ++i) {
      r.push_back(s[i]);
      if (s[i] == '\'' && i + 2 < s.size() && s[i + 1] == '\'') ++i;
    }
    return r;
  }
  if (s.size() >= 2 && s.front() == '"' && s.back() == '"') {
    std::string r;
    for (size_t i = 1; i + 1 < s.size(); ++i) {
      char c = s[i];
      if (c == '\\' && i + 2 < s.size()) {
        char n = s[++i];
        switch (n) {
          case 'n': r.push_back('\n'); break;
          case 't': r.push_back('\t'); break;
          case '\\': r.push_back('\\'); break;
          case '"': r.push_back('"'); break;
          default: r.push_back(n); break;
        }
      } else {
        r.push_back(c);
      }
    }
    return r;
  }
  if (!s.empty() && (s.front() == '\'' || s.front() == '"'))
    throw ParseError("unterminated quoted scalar at line " + std::to_string(lineno));
  return s;
}

Node scalarNode(const std::string& raw, int lineno) {
  Node n;
  std::string t = trim(raw);
  if (t.empty() || t == "~" || t == "null") return n;  // Null
  n.kind = Node::Scalar;
  n.text = unquote(t, lineno);
  return n;
}

// Position of the ':' that separates a mapping key from its value, or npos.
size_t findKeyColon(const std::string& s) {
  bool sq = false, dq = false;
  int depth = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (sq || dq) continue;
    else if (c == '[' || c == '{') ++depth;
    else if (c == ']' || c == '}') --depth;
    else if (c == ':' && depth == 0 && (i + 1 == s.size() || s[i + 1] == ' ')) return i;
  }
  return std::string::npos;
}

bool isSeqItem(const std::string& s) { return s == "-" || (s.size() >= 2 && s[0] == '-' && s[1] == ' '); }

bool bracketsBalanced(const std::string& s) {
  bool sq = false, dq = false;
  int depth = 0;
  for (char c : s) {
    if (c == '\'' && !dq) sq = !sq;
    else if (c == '"' && !sq) dq = !dq;
    else if (sq || dq) continue;
    else if (c == '[' || c == '{') ++depth;
    else if (c == ']' || c == '}') --depth;
  }
  return depth <= 0;
}

// ---- flow collections -------------------------------------------------------
struct FlowParser {
  const std::string& s;
  size_t i = 0;
  int lineno;
  FlowParser(const std::string& str, int ln) : s(str), lineno(ln) {}

  void ws() {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n')) ++i;
  }
  [[noreturn]] void fail(const char* what) {
    throw ParseError(std::string("flow collection: ") + what + " at line " + std::to_string(lineno));
  }
  std::string scalarText(const char* stops) {
    ws();
    size_t b = i;
    if (i < s.size() && (s[i] == '\'' || s[i] == '"')) {
      const char q = s[i++];
      while (i < s.size()) {
        if (s[i] == q) {
          if (q == '\'' && i + 1 < s.size() && s[i + 1] == '\'') { i += 2; continue; }
          break;
        }
        if (q == '"' && s[i] == '\\') ++i;
        ++i;
      }
      if (i >= s.size()) fail("unterminated quote");
      ++i;
      return s.substr(b, i - b);
    }
    while (i < s.size()) {
      bool stop = false;
      for (const char* p = stops; *p; ++p)
        if (s[i] == *p) stop = true;
      if (s[i] == ':' && (i + 1 == s.size() || s[i + 1] == ' ') && std::string(stops).find(':') != std::string::npos)
        stop = true;
      if (stop) break;
      ++i;
    }
    return s.substr(b, i - b);
  }
  Node value() {
    ws();
    if (i >= s.size()) fail("unexpected end");
    if (s[i] == '[') return seq();
    if (s[i] == '{') return map();
    return scalarNode(scalarText(",]}"), lineno);
  }
  Node seq() {
    Node n;
    n.kind = Node::Sequence;
    ++i;  // '['
    ws();
    if (i < s.size() && s[i] == ']') { ++i; return n; }
    while (true) {
      n.items.push_back(value());
      ws();
      if (i >= s.size()) fail("missing ]");
      if (s[i] == ',') { ++i; ws(); if (i < s.size() && s[i] == ']') { ++i; return n; } continue; }
      if (s[i] == ']') { ++i; return n; }
      fail("expected , or ]");
    }
  }
  Node map() {
    Node n;
    n.kind = Node::Map;
    ++i;  // '{'
    ws();
    if (i < s.size() && s[i] == '}') { ++i; return n; }
    while (true) {
      std::string key = unquote(trim(scalarText(":,}")), lineno);
      ws();
      Node v;
      if (i < s.size() && s[i] == ':') {
        ++i;
        ws();
        if (i < s.size() && (s[i] == ',' || s[i] == '}')) v = Node();
        else v = value();
      }
      n.entries.emplace_back(key, std::move(v));
      ws();
      if (i >= s.size()) fail("missing }");
      if (s[i] == ',') { ++i; continue; }
      if (s[i] == '}') { ++i; return n; }
      fail("expected , or }");
    }
  }
};

Node parseFlow(const std::string& s, int lineno) {
  FlowParser p(s, lineno);
  Node n = p.value();
  p.ws();
  if (p.i != s.size()) throw ParseError("trailing text after flow collection at line " + std::to_string(lineno));
  return n;
}

// ---- block structure -------------------------------------------------------
struct BlockParser {
  std::vector<Line> lines;
  size_t pos = 0;
  std::vector<std::pair<std::string, Node>> anchors;  // "&name" definitions seen so far

  // Splits a leading "&anchor" off `text` (returns the name, "" if none).
  static std::string takeAnchor(std::string& text) {
    if (text.empty() || text[0] != '&') return std::string();
    size_t e = text.find(' ');
    std::string name = text.substr(1, e == std::string::npos ? std::string::npos : e - 1);
    text = e == std::string::npos ? std::string() : trim(text.substr(e + 1));
    return name;
  }
  Node alias(const std::string& text, int lineno) {
    const std::string name = trim(text.substr(1));
    for (auto it = anchors.rbegin(); it != anchors.rend(); ++it)
      if (it->first == name) return it->second;
    throw ParseError("unknown alias *" + name + " at line " + std::to_string(lineno));
  }

  bool atEnd() const { return pos >= lines.size(); }

  // Value written inline after "key:" or "- ", possibly continued on the
  // following lines that are indented deeper than `parentIndent`.
  Node inlineValue(std::string text, int parentIndent, int lineno) {
    text = trim(text);
    if (!text.empty() && (text[0] == '[' || text[0] == '{')) {
      while (!bracketsBalanced(text) && !atEnd() && lines[pos].indent > parentIndent) {
        text += " " + lines[pos].text;
        ++pos;
      }
      return parseFlow(text, lineno);
    }
    if (!text.empty() && (text[0] == '|' || text[0] == '>'))
      throw ParseError("block scalars are not supported (line " + std::to_string(lineno) + ")");
    // multi-line plain scalar: join continuation lines with single spaces
    if (text.empty() || (text[0] != '\'' && text[0] != '"')) {
      while (!atEnd() && lines[pos].indent > parentIndent && findKeyColon(lines[pos].text) == std::string::npos &&
             !isSeqItem(lines[pos].text)) {
        text += " " + lines[pos].text;
        ++pos;
      }
    }
    return scalarNode(text, lineno);
  }

  Node block(int indent) {
    if (atEnd()) return Node();
    const Line& ln = lines[pos];
    if (ln.indent != indent) throw ParseError("bad indentation at line " + std::to_string(ln.lineno));
    if (isSeqItem(ln.text)) return sequence(indent);
    if (findKeyColon(ln.text) != std::string::npos) return mapping(indent);
    ++pos;
    return inlineValue(ln.text, indent - 1 < 0 ? -1 : indent - 1, ln.lineno);
  }

  Node sequence(int indent) {
    Node n;
    n.kind = Node::Sequence;
    while (!atEnd() && lines[pos].indent == indent && isSeqItem(lines[pos].text)) {
      Line& ln = lines[pos];
      size_t b = ln.text.find_first_not_of(' ', 1);
      if (b == std::string::npos) {  // "-" alone: the item is the nested block
        ++pos;
        if (!atEnd() && lines[pos].indent > indent) n.items.push_back(block(lines[pos].indent));
        else n.items.push_back(Node());
        continue;
      }
      const std::string rest = ln.text.substr(b);
      const int col = indent + int(b);
      if (isSeqItem(rest) || findKeyColon(rest) != std::string::npos) {
        // rewrite this line as a virtual line starting at the item's column
        ln.indent = col;
        ln.text = rest;
        n.items.push_back(block(col));
      } else if (rest[0] == '*') {
        ++pos;
        n.items.push_back(alias(rest, ln.lineno));
      } else {
        ++pos;
        n.items.push_back(inlineValue(rest, indent, ln.lineno));
      }
    }
    return n;
  }

  Node mapping(int indent) {
    Node n;
    n.kind = Node::Map;
    while (!atEnd() && lines[pos].indent == indent && !isSeqItem(lines[pos].text)) {
      const Line ln = lines[pos];
      const size_t c = findKeyColon(ln.text);
      if (c == std::string::npos) throw ParseError("expected 'key:' at line " + std::to_string(ln.lineno));
      std::string key = unquote(trim(ln.text.substr(0, c)), ln.lineno);
      std::string val = trim(ln.text.substr(c + 1));
      ++pos;
      const std::string anchor = takeAnchor(val);
      Node v;
      if (!val.empty() && val[0] == '*') {
        v = alias(val, ln.lineno);
      } else if (!val.empty()) {
        v = inlineValue(val, indent, ln.lineno);
      } else if (!atEnd() && lines[pos].indent > indent) {
        v = block(lines[pos].indent);
      } else if (!atEnd() && lines[pos].indent == indent && isSeqItem(lines[pos].text)) {
        v = sequence(indent);  // sequence written at the key's own column
      }
      if (!anchor.empty()) anchors.emplace_back(anchor, v);
      n.entries.emplace_back(std::move(key), std::move(v));
    }
    return n;
  }
};

}  // namespace

const Node* Node::find(const std::string& key) const {
  if (kind != Map) return nullptr;
  for (const auto& e : entries)
    if (e.first == key) return &e.second;
  return nullptr;
}

const Node& Node::operator[](const std::string& key) const {
  const Node* n = find(key);
  if (!n) throw ParseError("missing key '" + key + "'");
  return *n;
}

const std::string& Node::str() const {
  if (kind != Scalar) throw ParseError("node is not a scalar");
  return text;
}

double Node::asDouble() const {
  const std::string& s = str();
  errno = 0;
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end != '\0' || (errno == ERANGE && v != 0.0))
    throw ParseError("not a number: '" + s + "'");
  return v;
}

int Node::asInt() const {
  const std::string& s = str();
  char* end = nullptr;
  const long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str() || *end != '\0') throw ParseError("not an integer: '" + s + "'");
  return int(v);
}

Node parse(const std::string& text) {
  BlockParser p;
  p.lines = splitLines(text);
  if (p.lines.empty()) return Node();
  Node root = p.block(p.lines[0].indent);
  if (!p.atEnd())
    throw ParseError("unexpected content at line " + std::to_string(p.lines[p.pos].lineno));
  return root;
}

Node parseFile(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw ParseError("cannot open '" + path + "'");
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

}  // namespace yaml
}  // namespace tchem_amd
