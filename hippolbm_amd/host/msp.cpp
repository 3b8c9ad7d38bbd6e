// msp.cpp -- the .msp (YAML subset) reader; see msp.hpp.
#include "msp.hpp"

#include <cctype>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace msp {

namespace {

struct Line {
  int indent;
  std::string text;
  int lineno;
};

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}

// drop a comment: '#' at the start or after whitespace, outside quotes
std::string strip_comment(const std::string& s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); i++) {
    const char c = s[i];
    if (q) {
      if (c == q) q = 0;
    } else if (c == '"' || c == '\'') {
      q = c;
    } else if (c == '#' && (i == 0 || std::isspace((unsigned char)s[i - 1]))) {
      return s.substr(0, i);
    }
  }
  return s;
}

bool is_seq_item(const std::string& t) { return t == "-" || (t.size() >= 2 && t[0] == '-' && t[1] == ' '); }

// position of the key separator ':' (followed by a blank or the end), or npos
size_t key_sep(const std::string& t) {
  if (t.empty() || t[0] == '[' || t[0] == '{' || t[0] == '"' || t[0] == '\'') return std::string::npos;
  for (size_t i = 0; i < t.size(); i++) {
    const char c = t[i];
    if (c == '[' || c == '{') return std::string::npos;
    if (c == ':' && (i + 1 == t.size() || t[i + 1] == ' ')) return i;
  }
  return std::string::npos;
}

struct FlowParser {
  const std::string& s;
  size_t i = 0;
  int lineno;
  explicit FlowParser(const std::string& str, int ln) : s(str), lineno(ln) {}
  void ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
  }
  [[noreturn]] void fail(const std::string& m) {
    throw ParseError("line " + std::to_string(lineno) + ": " + m + " in '" + s + "'");
  }
  Node value() {
    ws();
    if (i >= s.size()) fail("missing value");
    if (s[i] == '[') return seq();
    if (s[i] == '{') return map();
    return scalar("],}");
  }
  Node scalar(const char* stops) {
    ws();
    Node n;
    n.kind = Node::SCALAR;
    if (i < s.size() && (s[i] == '"' || s[i] == '\'')) {
      const char q = s[i++];
      const size_t e = s.find(q, i);
      if (e == std::string::npos) fail("unterminated quote");
      n.text = s.substr(i, e - i);
      i = e + 1;
      return n;
    }
    const size_t b = i;
    while (i < s.size() && s[i] != ',' && !std::strchr(stops, s[i])) i++;
    n.text = trim(s.substr(b, i - b));
    if (n.text == "~" || n.text == "null") n.kind = Node::NUL;
    return n;
  }
  Node seq() {
    Node n;
    n.kind = Node::SEQ;
    i++;  // '['
    ws();
    if (i < s.size() && s[i] == ']') {
      i++;
      return n;
    }
    for (;;) {
      n.items.push_back(value());
      ws();
      if (i >= s.size()) fail("unterminated [");
      if (s[i] == ',') {
        i++;
        continue;
      }
      if (s[i] == ']') {
        i++;
        return n;
      }
      fail("expected , or ]");
    }
  }
  Node map() {
    Node n;
    n.kind = Node::MAP;
    i++;  // '{'
    ws();
    if (i < s.size() && s[i] == '}') {
      i++;
      return n;
    }
    for (;;) {
      ws();
      const size_t b = i;
      while (i < s.size() && s[i] != ':') i++;
      if (i >= s.size()) fail("expected key:");
      const std::string key = trim(s.substr(b, i - b));
      i++;
      n.fields.emplace_back(key, value());
      ws();
      if (i >= s.size()) fail("unterminated {");
      if (s[i] == ',') {
        i++;
        continue;
      }
      if (s[i] == '}') {
        i++;
        return n;
      }
      fail("expected , or }");
    }
  }
};

Node inline_value(const std::string& t, int lineno) {
  FlowParser p(t, lineno);
  Node n = p.value();
  p.ws();
  if (p.i != t.size()) {
    // a plain scalar may contain ',' ']' '}' outside flow context: take the whole text
    if (t[0] != '[' && t[0] != '{') {
      Node s;
      s.kind = Node::SCALAR;
      s.text = t;
      return s;
    }
    p.fail("trailing characters");
  }
  return n;
}

struct BlockParser {
  std::vector<Line> L;
  size_t pos = 0;

  [[noreturn]] void fail(const std::string& m) {
    const int ln = pos < L.size() ? L[pos].lineno : (L.empty() ? 0 : L.back().lineno);
    throw ParseError("line " + std::to_string(ln) + ": " + m);
  }

  Node block(int indent) {
    if (pos >= L.size() || L[pos].indent < indent) return Node{};
    if (is_seq_item(L[pos].text)) return seq(L[pos].indent);
    return map(L[pos].indent);
  }

  Node seq(int ind) {
    Node n;
    n.kind = Node::SEQ;
    while (pos < L.size() && L[pos].indent == ind && is_seq_item(L[pos].text)) {
      const std::string& t = L[pos].text;
      size_t off = 1;
      while (off < t.size() && t[off] == ' ') off++;
      const std::string rest = t.substr(off);
      if (rest.empty()) {
        pos++;
        n.items.push_back(pos < L.size() && L[pos].indent > ind ? block(L[pos].indent) : Node{});
      } else if (key_sep(rest) != std::string::npos) {
        // "- key: ..." opens a map whose keys sit at the column of `key`
        L[pos].indent = ind + (int)off;
        L[pos].text = rest;
        n.items.push_back(map(ind + (int)off));
      } else {
        n.items.push_back(inline_value(rest, L[pos].lineno));
        pos++;
      }
    }
    return n;
  }

  Node map(int ind) {
    Node n;
    n.kind = Node::MAP;
    while (pos < L.size() && L[pos].indent == ind && !is_seq_item(L[pos].text)) {
      const std::string t = L[pos].text;
      const size_t c = key_sep(t);
      if (c == std::string::npos) fail("expected 'key: value', got '" + t + "'");
      const std::string key = trim(t.substr(0, c));
      const std::string rest = trim(t.substr(c + 1));
      const int ln = L[pos].lineno;
      pos++;
      Node v;
      if (!rest.empty() && (rest[0] == '|' || rest[0] == '>')) {
        v.kind = Node::SCALAR;
        while (pos < L.size() && L[pos].indent > ind) {
          if (!v.text.empty()) v.text += '\n';
          v.text += L[pos].text;
          pos++;
        }
      } else if (!rest.empty()) {
        v = inline_value(rest, ln);
      } else if (pos < L.size() && L[pos].indent > ind) {
        v = block(L[pos].indent);
      } else if (pos < L.size() && L[pos].indent == ind && is_seq_item(L[pos].text)) {
        v = seq(ind);
      }
      n.fields.emplace_back(key, std::move(v));
    }
    return n;
  }
};

std::string json_escape(const std::string& s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') {
      o += '\\';
      o += c;
    } else if (c == '\n') {
      o += "\\n";
    } else if ((unsigned char)c < 0x20) {
      o += ' ';
    } else {
      o += c;
    }
  }
  return o;
}

}  // namespace

const Node* Node::find(const std::string& key) const {
  if (kind != MAP) return nullptr;
  for (const auto& f : fields)
    if (f.first == key) return &f.second;
  return nullptr;
}

const Node& Node::at(const std::string& key) const {
  const Node* n = find(key);
  if (!n) throw ParseError("missing key '" + key + "'");
  return *n;
}

double Node::number() const {
  if (kind != SCALAR) throw ParseError("expected a number");
  const char* b = text.c_str();
  char* e = nullptr;
  const double v = std::strtod(b, &e);
  if (e == b) throw ParseError("not a number: '" + text + "'");
  return v;  // a trailing unit ("0.01 m", "1e-5 m2.s-1") is ignored
}

long Node::integer() const {
  const double v = number();
  return (long)v;
}

bool Node::boolean() const {
  if (kind != SCALAR) throw ParseError("expected a boolean");
  if (text == "true" || text == "True" || text == "yes" || text == "on") return true;
  if (text == "false" || text == "False" || text == "no" || text == "off") return false;
  throw ParseError("not a boolean: '" + text + "'");
}

std::vector<double> Node::numbers() const {
  if (kind != SEQ) throw ParseError("expected a sequence of numbers");
  std::vector<double> v;
  for (const auto& it : items) v.push_back(it.number());
  return v;
}

Node parse(const std::string& text) {
  BlockParser p;
  std::istringstream in(text);
  std::string raw;
  int ln = 0;
  bool in_block_scalar = false;
  int block_indent = 0;
  while (std::getline(in, raw)) {
    ln++;
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();
    int ind = 0;
    while (ind < (int)raw.size() && raw[ind] == ' ') ind++;
    if (ind < (int)raw.size() && raw[ind] == '\t') throw ParseError("line " + std::to_string(ln) + ": tab indentation");
    if (in_block_scalar) {
      if (trim(raw).empty()) continue;
      if (ind > block_indent) {  // block scalar text is kept verbatim (no comment stripping)
        p.L.push_back({ind, trim(raw), ln});
        continue;
      }
      in_block_scalar = false;
    }
    const std::string t = trim(strip_comment(raw));
    if (t.empty()) continue;
    p.L.push_back({ind, t, ln});
    const size_t c = key_sep(t);
    if (c != std::string::npos) {
      const std::string rest = trim(t.substr(c + 1));
      if (!rest.empty() && (rest[0] == '|' || rest[0] == '>')) {
        in_block_scalar = true;
        block_indent = ind;
      }
    }
  }
  if (p.L.empty()) return Node{};
  Node root = p.block(p.L[0].indent);
  if (p.pos != p.L.size()) p.fail("unexpected indentation");
  return root;
}

Node parse_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw ParseError("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str());
}

namespace {

void merge_top(Node& into, const Node& from) {
  for (const auto& kv : from.fields) {
    if (kv.first == "includes") continue;
    bool replaced = false;
    for (auto& mine : into.fields)
      if (mine.first == kv.first) {
        mine.second = kv.second;
        replaced = true;
      }
    if (!replaced) into.fields.push_back(kv);
  }
}

Node load_rec(const std::string& path, int depth) {
  if (depth > 16) throw ParseError(path + ": includes nested deeper than 16 files");
  Node doc = parse_file(path);
  if (doc.is_null()) return doc;
  if (!doc.is_map()) throw ParseError(path + ": a deck is a map of blocks");
  const Node* inc = doc.find("includes");
  if (!inc || inc->is_null()) return doc;
  const size_t slash = path.find_last_of('/');
  const std::string dir = slash == std::string::npos ? std::string() : path.substr(0, slash + 1);
  Node merged;
  merged.kind = Node::MAP;
  auto one = [&](const Node& item) {
    if (!item.is_scalar()) throw ParseError(path + ": includes lists file names");
    const std::string p = !item.text.empty() && item.text[0] == '/' ? item.text : dir + item.text;
    std::ifstream probe(p);
    if (!probe) throw ParseError(path + ": included file '" + item.text + "' not found (looked for " + p + ")");
    merge_top(merged, load_rec(p, depth + 1));
  };
  if (inc->is_seq())
    for (const auto& item : inc->items) one(item);
  else
    one(*inc);
  merge_top(merged, doc);
  return merged;
}

}  // namespace

Node load_deck(const std::string& path) { return load_rec(path, 0); }

std::string to_json(const Node& n) {
  switch (n.kind) {
    case Node::NUL:
      return "null";
    case Node::SCALAR:
      return "\"" + json_escape(n.text) + "\"";
    case Node::SEQ: {
      std::string o = "[";
      for (size_t i = 0; i < n.items.size(); i++) o += (i ? "," : "") + to_json(n.items[i]);
      return o + "]";
    }
    default: {
      std::string o = "{";
      for (size_t i = 0; i < n.fields.size(); i++)
        o += (i ? ",\"" : "\"") + json_escape(n.fields[i].first) + "\":" + to_json(n.fields[i].second);
      return o + "}";
    }
  }
}

}  // namespace msp
