// frontend.cpp -- model files in the reference's YAML format -> bases and
// non-branching terms, with no Haskell in the loop.
//
// Reference behaviour followed:
//   * configuration layout: basis + hamiltonian (+ observables), Yaml.hs:14-48;
//     basis header keys and checks, Basis.hs:261-345;
//   * expression grammar (sums of optionally scaled products of primitives,
//     parentheses, unicode or ASCII spellings of sigma / S / c / c-dagger / n
//     with subscript indices and spin arrows), Parser.hs:60-170;
//   * particle type of an expression, Expr.hs:360-368;
//   * "sites" replication: the expression's distinct site indices, sorted,
//     are replaced by every listed tuple and the copies summed,
//     Expr.hs:266-286;
//   * generators as non-branching terms and their products,
//     Generator.hs:229-247, NonbranchingTerm.hs:43-58; sigma^x = s+ + s-,
//     sigma^y = -i s+ + i s-, S = sigma / 2 (Parser.hs:110-125 semantics).
//
// The YAML reader covers the subset the reference's model files use: block
// mappings and sequences by indentation, "- key: value" items, flow
// sequences (also over several lines), anchors and aliases, quoted and plain
// scalars, comments.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "operator_internal.hpp"

namespace lsamd {
namespace {

// ---------------------------------------------------------------------------
// YAML subset
// ---------------------------------------------------------------------------

struct Node;
using NodeP = std::shared_ptr<Node>;
struct Node {
  enum Kind { kNull, kScalar, kSeq, kMap } kind = kNull;
  std::string text;
  bool quoted = false;
  std::vector<NodeP> items;
  std::vector<std::pair<std::string, NodeP>> entries;
  NodeP get(std::string const &key) const {
    for (auto const &e : entries)
      if (e.first == key) return e.second;
    return nullptr;
  }
};

struct YamlError {
  std::string msg;
};

struct Line {
  int indent;
  std::string text; // without indentation and comment
  int number;
};

std::string rstrip(std::string s) {
  while (!s.empty() && std::isspace(static_cast<unsigned char>(s.back()))) s.pop_back();
  return s;
}

std::string strip(std::string const &s) {
  size_t b = 0;
  while (b < s.size() && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
  return rstrip(s.substr(b));
}

// Removes a trailing comment: '#' at the start or after whitespace, outside quotes.
std::string drop_comment(std::string const &s) {
  char q = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    char const c = s[i];
    if (q != 0) {
      if (c == q) q = 0;
      continue;
    }
    if (c == '"' || c == '\'') {
      q = c;
    } else if (c == '#' && (i == 0 || std::isspace(static_cast<unsigned char>(s[i - 1])))) {
      return s.substr(0, i);
    }
  }
  return s;
}

int bracket_balance(std::string const &s) {
  int d = 0;
  char q = 0;
  for (char c : s) {
    if (q != 0) {
      if (c == q) q = 0;
      continue;
    }
    if (c == '"' || c == '\'') q = c;
    else if (c == '[' || c == '{') ++d;
    else if (c == ']' || c == '}') --d;
  }
  return d;
}

std::vector<Line> split_lines(std::string const &text) {
  std::vector<Line> out;
  size_t pos = 0;
  int number = 0;
  while (pos <= text.size()) {
    size_t e = text.find('\n', pos);
    if (e == std::string::npos) e = text.size();
    std::string raw = text.substr(pos, e - pos);
    ++number;
    pos = e + 1;
    if (!raw.empty() && raw.back() == '\r') raw.pop_back();
    if (raw.find('\t') != std::string::npos && strip(drop_comment(raw)).size() > 0) {
      size_t k = 0;
      while (k < raw.size() && raw[k] == ' ') ++k;
      if (k < raw.size() && raw[k] == '\t') throw YamlError{"tab indentation at line " + std::to_string(number)};
    }
    std::string body = rstrip(drop_comment(raw));
    int ind = 0;
    while (ind < static_cast<int>(body.size()) && body[ind] == ' ') ++ind;
    if (ind == static_cast<int>(body.size())) continue;
    if (body.rfind("---", 0) == 0 || body.rfind("...", 0) == 0) continue;
    out.push_back(Line{ind, body.substr(ind), number});
    if (e == text.size()) break;
  }
  // join flow collections that span several lines
  std::vector<Line> joined;
  for (size_t i = 0; i < out.size(); ++i) {
    Line l = out[i];
    int bal = bracket_balance(l.text);
    while (bal > 0 && i + 1 < out.size()) {
      ++i;
      l.text += " " + out[i].text;
      bal = bracket_balance(l.text);
    }
    if (bal != 0) throw YamlError{"unbalanced brackets at line " + std::to_string(l.number)};
    joined.push_back(l);
  }
  return joined;
}

struct Yaml {
  std::vector<Line> lines;
  std::map<std::string, NodeP> anchors;

  static NodeP scalar(std::string s, bool quoted) {
    auto n = std::make_shared<Node>();
    n->kind = Node::kScalar;
    n->text = std::move(s);
    n->quoted = quoted;
    return n;
  }

  // Inline value: anchor / alias / flow collection / quoted / plain scalar.
  NodeP parse_inline(std::string const &s0, int line) {
    std::string s = strip(s0);
    if (s.empty()) return std::make_shared<Node>();
    size_t pos = 0;
    NodeP n = parse_flow(s, pos, line);
    while (pos < s.size() && std::isspace(static_cast<unsigned char>(s[pos]))) ++pos;
    if (pos != s.size()) throw YamlError{"unexpected text after a value at line " + std::to_string(line)};
    return n;
  }

  NodeP parse_flow(std::string const &s, size_t &pos, int line) {
    auto skip = [&] {
      while (pos < s.size() && std::isspace(static_cast<unsigned char>(s[pos]))) ++pos;
    };
    skip();
    if (pos >= s.size()) return std::make_shared<Node>();
    char const c = s[pos];
    if (c == '&') {
      size_t e = pos + 1;
      while (e < s.size() && !std::isspace(static_cast<unsigned char>(s[e])) && s[e] != ',' && s[e] != ']') ++e;
      std::string const name = s.substr(pos + 1, e - pos - 1);
      pos = e;
      NodeP v = parse_flow(s, pos, line);
      anchors[name] = v;
      return v;
    }
    if (c == '*') {
      size_t e = pos + 1;
      while (e < s.size() && !std::isspace(static_cast<unsigned char>(s[e])) && s[e] != ',' && s[e] != ']' &&
             s[e] != '}')
        ++e;
      std::string const name = s.substr(pos + 1, e - pos - 1);
      pos = e;
      auto it = anchors.find(name);
      if (it == anchors.end()) throw YamlError{"unknown alias *" + name + " at line " + std::to_string(line)};
      return it->second;
    }
    if (c == '[') {
      auto n = std::make_shared<Node>();
      n->kind = Node::kSeq;
      ++pos;
      skip();
      if (pos < s.size() && s[pos] == ']') {
        ++pos;
        return n;
      }
      while (true) {
        n->items.push_back(parse_flow(s, pos, line));
        skip();
        if (pos >= s.size()) throw YamlError{"unterminated [ at line " + std::to_string(line)};
        if (s[pos] == ',') {
          ++pos;
          skip();
          if (pos < s.size() && s[pos] == ']') { // trailing comma
            ++pos;
            return n;
          }
          continue;
        }
        if (s[pos] == ']') {
          ++pos;
          return n;
        }
        throw YamlError{"expected , or ] at line " + std::to_string(line)};
      }
    }
    if (c == '{') {
      auto n = std::make_shared<Node>();
      n->kind = Node::kMap;
      ++pos;
      skip();
      if (pos < s.size() && s[pos] == '}') {
        ++pos;
        return n;
      }
      while (true) {
        size_t e = pos;
        while (e < s.size() && s[e] != ':') ++e;
        if (e >= s.size()) throw YamlError{"expected key: in { } at line " + std::to_string(line)};
        std::string key = strip(s.substr(pos, e - pos));
        if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'')) key = key.substr(1, key.size() - 2);
        pos = e + 1;
        n->entries.emplace_back(key, parse_flow(s, pos, line));
        skip();
        if (pos < s.size() && s[pos] == ',') {
          ++pos;
          skip();
          continue;
        }
        if (pos < s.size() && s[pos] == '}') {
          ++pos;
          return n;
        }
        throw YamlError{"expected , or } at line " + std::to_string(line)};
      }
    }
    if (c == '"' || c == '\'') {
      std::string out;
      size_t i = pos + 1;
      for (; i < s.size(); ++i) {
        if (s[i] == c) {
          if (c == '\'' && i + 1 < s.size() && s[i + 1] == '\'') {
            out += '\'';
            ++i;
            continue;
          }
          break;
        }
        if (c == '"' && s[i] == '\\' && i + 1 < s.size()) {
          char const esc = s[++i];
          switch (esc) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case '"': out += '"'; break;
          case '\\': out += '\\'; break;
          default: out += '\\'; out += esc; break;
          }
          continue;
        }
        out += s[i];
      }
      if (i >= s.size()) throw YamlError{"unterminated string at line " + std::to_string(line)};
      pos = i + 1;
      return scalar(out, true);
    }
    // plain scalar: up to , ] } in flow context (a plain top-level scalar has none of them)
    size_t e = pos;
    while (e < s.size() && s[e] != ',' && s[e] != ']' && s[e] != '}') ++e;
    std::string const v = strip(s.substr(pos, e - pos));
    pos = e;
    if (v == "~" || v == "null") return std::make_shared<Node>();
    return scalar(v, false);
  }

  // "key: rest" split; returns false if the line is not a mapping entry.
  static bool split_key(std::string const &t, std::string &key, std::string &rest) {
    char q = 0;
    int depth = 0;
    for (size_t i = 0; i < t.size(); ++i) {
      char const c = t[i];
      if (q != 0) {
        if (c == q) q = 0;
        continue;
      }
      if (c == '"' || c == '\'') {
        if (i == 0) q = c;
        continue;
      }
      if (c == '[' || c == '{') ++depth;
      if (c == ']' || c == '}') --depth;
      if (c == ':' && depth == 0 && (i + 1 == t.size() || t[i + 1] == ' ')) {
        key = strip(t.substr(0, i));
        if (key.size() >= 2 && (key[0] == '"' || key[0] == '\'') && key.back() == key[0])
          key = key.substr(1, key.size() - 2);
        rest = i + 1 < t.size() ? strip(t.substr(i + 1)) : std::string();
        return !key.empty();
      }
    }
    return false;
  }

  static bool is_item(std::string const &t) { return t == "-" || t.rfind("- ", 0) == 0; }

  NodeP parse_block(size_t &i, int indent) {
    if (i >= lines.size()) return std::make_shared<Node>();
    Line const &first = lines[i];
    if (is_item(first.text)) {
      auto n = std::make_shared<Node>();
      n->kind = Node::kSeq;
      while (i < lines.size() && lines[i].indent == indent && is_item(lines[i].text)) {
        std::string const rest = lines[i].text == "-" ? std::string() : strip(lines[i].text.substr(2));
        if (rest.empty()) {
          ++i;
          if (i < lines.size() && lines[i].indent > indent) {
            n->items.push_back(parse_block(i, lines[i].indent));
          } else {
            n->items.push_back(std::make_shared<Node>());
          }
          continue;
        }
        std::string key, val;
        if (split_key(rest, key, val) || is_item(rest)) {
          // the item is a block collection starting on this line
          int const col = indent + static_cast<int>(lines[i].text.size() - rest.size());
          lines[i].indent = col;
          lines[i].text = rest;
          n->items.push_back(parse_block(i, col));
        } else {
          n->items.push_back(parse_inline(rest, lines[i].number));
          ++i;
        }
      }
      if (i < lines.size() && lines[i].indent > indent)
        throw YamlError{"bad indentation at line " + std::to_string(lines[i].number)};
      return n;
    }
    auto n = std::make_shared<Node>();
    n->kind = Node::kMap;
    while (i < lines.size() && lines[i].indent == indent && !is_item(lines[i].text)) {
      std::string key, rest;
      if (!split_key(lines[i].text, key, rest))
        throw YamlError{"expected 'key: value' at line " + std::to_string(lines[i].number)};
      int const line = lines[i].number;
      ++i;
      std::string anchor;
      if (!rest.empty() && rest[0] == '&' && rest.find(' ') == std::string::npos) {
        anchor = rest.substr(1);
        rest.clear();
      }
      NodeP v;
      if (rest.empty()) {
        if (i < lines.size() && (lines[i].indent > indent || (lines[i].indent == indent && is_item(lines[i].text)))) {
          v = parse_block(i, lines[i].indent);
        } else {
          v = std::make_shared<Node>();
        }
      } else {
        v = parse_inline(rest, line);
      }
      if (!anchor.empty()) anchors[anchor] = v;
      for (auto const &e : n->entries)
        if (e.first == key) throw YamlError{"duplicate key '" + key + "' at line " + std::to_string(line)};
      n->entries.emplace_back(key, v);
    }
    if (i < lines.size() && lines[i].indent > indent)
      throw YamlError{"bad indentation at line " + std::to_string(lines[i].number)};
    return n;
  }

  NodeP parse(std::string const &text) {
    lines = split_lines(text);
    size_t i = 0;
    if (lines.empty()) return std::make_shared<Node>();
    NodeP root = parse_block(i, lines[0].indent);
    if (i != lines.size()) throw YamlError{"unexpected content at line " + std::to_string(lines[i].number)};
    return root;
  }
};

// ---------------------------------------------------------------------------
// Non-branching term algebra
// ---------------------------------------------------------------------------

using cplx = std::complex<double>;
struct Key {
  uint64_t m, l, r, x, s;
  bool operator<(Key const &o) const { return std::tie(m, l, r, x, s) < std::tie(o.m, o.l, o.r, o.x, o.s); }
};
using Poly = std::map<Key, cplx>;

void add_to(Poly &p, Key const &k, cplx v) {
  auto it = p.find(k);
  if (it == p.end()) {
    if (v != cplx(0.0)) p.emplace(k, v);
  } else {
    it->second += v;
    if (it->second == cplx(0.0)) p.erase(it);
  }
}

Poly poly_add(Poly a, Poly const &b) {
  for (auto const &[k, v] : b) add_to(a, k, v);
  return a;
}

Poly poly_scale(Poly const &a, cplx c) {
  Poly out;
  if (c == cplx(0.0)) return out;
  for (auto const &[k, v] : a) out.emplace(k, v * c);
  return out;
}

// Product a * b (b acts first): the composition of two non-branching terms
// (NonbranchingTerm.hs:43-58) with the sign of the second term's string
// evaluated on the intermediate state.
Poly poly_mul(Poly const &A, Poly const &B) {
  Poly out;
  for (auto const &[ka, va] : A)
    for (auto const &[kb, vb] : B) {
      if (((ka.r ^ kb.l) & ka.m & kb.m) != 0) continue;
      Key k;
      k.m = ka.m | kb.m;
      k.r = kb.r | (ka.r & ~kb.m);
      k.l = ka.l | (kb.l & ~ka.m);
      k.x = k.l ^ k.r;
      k.s = (ka.s ^ kb.s) & ~k.m;
      int const p = __builtin_popcountll(k.r & (ka.s ^ kb.s)) + __builtin_popcountll(kb.x & ka.s);
      cplx const v = ((p & 1) ? -1.0 : 1.0) * va * vb;
      add_to(out, k, v);
    }
  return out;
}

Poly term(cplx v, uint64_t m, uint64_t l, uint64_t r, uint64_t s) {
  Poly p;
  add_to(p, Key{m, l, r, l ^ r, s}, v);
  return p;
}

// Generators (Generator.hs:229-247).
Poly spin_z(int i) { return term(-1.0, 0, 0, 0, uint64_t{1} << i); }
Poly spin_plus(int i) {
  uint64_t const b = uint64_t{1} << i;
  return term(1.0, b, b, 0, 0);
}
Poly spin_minus(int i) {
  uint64_t const b = uint64_t{1} << i;
  return term(1.0, b, 0, b, 0);
}
Poly fermion_number(int i) {
  uint64_t const b = uint64_t{1} << i;
  return term(1.0, b, b, b, 0);
}
Poly fermion_create(int i) {
  uint64_t const b = uint64_t{1} << i;
  return term(1.0, b, b, 0, b - 1);
}
Poly fermion_annihilate(int i) {
  uint64_t const b = uint64_t{1} << i;
  return term(1.0, b, 0, b, b - 1);
}

// ---------------------------------------------------------------------------
// Expressions
// ---------------------------------------------------------------------------

enum Particle { kSpin = 0, kSpinlessFermion = 1, kSpinfulFermion = 2 };

struct SExpr {
  enum Kind { kSum, kScaled, kProduct, kSpinOp, kFermionOp } kind;
  std::vector<SExpr> children;
  double coeff = 1.0;
  char prefix = 's'; // 's' sigma, 'S' spin-1/2 operator
  char op = 0;       // spin: x y z + -; fermion: c (create) a (annihilate) n
  int site = 0;
  int spin = -1; // fermions: -1 none, 0 up, 1 down
};

struct ExprError {
  std::string msg;
};

struct ExprParser {
  std::string s;
  size_t pos = 0;

  bool at(char const *lit) const { return s.compare(pos, std::strlen(lit), lit) == 0; }
  bool eat(char const *lit) {
    if (at(lit)) {
      pos += std::strlen(lit);
      return true;
    }
    return false;
  }
  void space() {
    while (pos < s.size() && std::isspace(static_cast<unsigned char>(s[pos]))) ++pos;
  }
  [[noreturn]] void error(std::string const &what) const {
    throw ExprError{what + " at byte " + std::to_string(pos) + " of \"" + s + "\""};
  }

  bool subscript_digit(int &d) {
    static char const *const digits[] = {"₀", "₁", "₂", "₃", "₄", "₅", "₆", "₇", "₈", "₉"};
    for (int k = 0; k < 10; ++k)
      if (eat(digits[k])) {
        d = k;
        return true;
      }
    return false;
  }
  bool index(int &out) {
    int d;
    if (subscript_digit(d)) {
      long v = d;
      while (subscript_digit(d)) v = v * 10 + d;
      out = static_cast<int>(v);
      return true;
    }
    size_t const save = pos;
    eat("_");
    if (pos < s.size() && std::isdigit(static_cast<unsigned char>(s[pos]))) {
      long v = 0;
      while (pos < s.size() && std::isdigit(static_cast<unsigned char>(s[pos]))) v = v * 10 + (s[pos++] - '0');
      out = static_cast<int>(v);
      return true;
    }
    pos = save;
    return false;
  }
  bool spin_label(int &out) {
    if (eat("↑")) return out = 0, true;
    if (eat("↓")) return out = 1, true;
    size_t const save = pos;
    eat("_");
    if (eat("\\up")) return out = 0, true;
    if (eat("\\down")) return out = 1, true;
    pos = save;
    return false;
  }

  bool primitive(SExpr &e) {
    size_t const save = pos;
    // spin operators
    char prefix = 0;
    if (eat("σ") || eat("\\sigma")) prefix = 's';
    else if (eat("S")) prefix = 'S';
    if (prefix != 0) {
      char op = 0;
      if (eat("ˣ")) op = 'x';
      else if (eat("ʸ")) op = 'y';
      else if (eat("ᶻ")) op = 'z';
      else if (eat("⁺")) op = '+';
      else if (eat("⁻")) op = '-';
      else {
        eat("^");
        if (eat("x")) op = 'x';
        else if (eat("y")) op = 'y';
        else if (eat("z")) op = 'z';
        else if (eat("+")) op = '+';
        else if (eat("-")) op = '-';
      }
      int site;
      if (op == 0 || !index(site)) {
        pos = save;
        return false;
      }
      e = SExpr{SExpr::kSpinOp, {}};
      e.prefix = prefix;
      e.op = op;
      e.site = site;
      return true;
    }
    char op = 0;
    if (eat("c†") || eat("c^\\dagger") || eat("c\\dagger")) op = 'c';
    else if (eat("c")) op = 'a';
    else if (eat("n")) op = 'n';
    if (op == 0) return false;
    int spin = -1, site;
    spin_label(spin);
    if (!index(site)) {
      pos = save;
      return false;
    }
    if (spin < 0) spin_label(spin);
    e = SExpr{SExpr::kFermionOp, {}};
    e.op = op;
    e.site = site;
    e.spin = spin;
    return true;
  }

  bool number(double &out) {
    size_t e = pos;
    if (e < s.size() && (s[e] == '+' || s[e] == '-')) ++e;
    size_t const digits0 = e;
    while (e < s.size() && std::isdigit(static_cast<unsigned char>(s[e]))) ++e;
    if (e < s.size() && s[e] == '.') {
      ++e;
      while (e < s.size() && std::isdigit(static_cast<unsigned char>(s[e]))) ++e;
    }
    if (e == digits0 || (e == digits0 + 1 && s[digits0] == '.')) return false;
    if (e < s.size() && (s[e] == 'e' || s[e] == 'E')) {
      size_t f = e + 1;
      if (f < s.size() && (s[f] == '+' || s[f] == '-')) ++f;
      if (f < s.size() && std::isdigit(static_cast<unsigned char>(s[f]))) {
        while (f < s.size() && std::isdigit(static_cast<unsigned char>(s[f]))) ++f;
        e = f;
      }
    }
    out = std::strtod(s.substr(pos, e - pos).c_str(), nullptr);
    pos = e;
    return true;
  }

  SExpr product() {
    SExpr p{SExpr::kProduct, {}};
    while (true) {
      space();
      if (pos < s.size() && s[pos] == '(') {
        ++pos;
        SExpr inner = sum();
        space();
        if (!eat(")")) error("expected )");
        p.children.push_back(std::move(inner));
        continue;
      }
      SExpr prim{SExpr::kProduct, {}};
      if (primitive(prim)) {
        p.children.push_back(std::move(prim));
        continue;
      }
      break;
    }
    if (p.children.empty()) error("expected an operator");
    return p;
  }

  SExpr scaled() {
    space();
    double c = 1.0;
    size_t const save = pos;
    double v;
    if (number(v)) {
      space();
      eat("×");
      eat("*");
      c = v;
    } else {
      pos = save;
    }
    SExpr e{SExpr::kScaled, {}};
    e.coeff = c;
    e.children.push_back(product());
    return e;
  }

  SExpr sum() {
    SExpr e{SExpr::kSum, {}};
    while (true) {
      space();
      if (pos >= s.size() || s[pos] == ')') break;
      double sign = 1.0;
      if (!e.children.empty() || at("-") || at("+")) {
        if (eat("-")) sign = -1.0;
        else if (eat("+")) sign = 1.0;
        else error("expected + or -");
      }
      SExpr t = scaled();
      t.coeff *= sign;
      e.children.push_back(std::move(t));
    }
    return e;
  }

  SExpr parse(std::string const &text) {
    s = text;
    pos = 0;
    SExpr e = sum();
    space();
    if (pos != s.size()) error("unexpected character");
    return e;
  }
};

void collect_sites(SExpr const &e, std::set<int> &out) {
  if (e.kind == SExpr::kSpinOp || e.kind == SExpr::kFermionOp) out.insert(e.site);
  for (auto const &c : e.children) collect_sites(c, out);
}

int particle_of(SExpr const &e) {
  if (e.kind == SExpr::kSpinOp) return kSpin;
  if (e.kind == SExpr::kFermionOp) return e.spin >= 0 ? kSpinfulFermion : kSpinlessFermion;
  for (auto const &c : e.children) {
    int const p = particle_of(c);
    if (p >= 0) return p;
  }
  return -1;
}

struct Lowering {
  int particle;
  int number_sites;
  std::map<int, int> remap; // site replication
  int site(int i) const {
    auto it = remap.find(i);
    return it == remap.end() ? i : it->second;
  }
  Poly lower(SExpr const &e) const {
    switch (e.kind) {
    case SExpr::kSum: {
      Poly out;
      for (auto const &c : e.children) out = poly_add(std::move(out), lower(c));
      return out;
    }
    case SExpr::kScaled: return poly_scale(lower(e.children[0]), e.coeff);
    case SExpr::kProduct: {
      Poly out = lower(e.children[0]);
      for (size_t k = 1; k < e.children.size(); ++k) out = poly_mul(out, lower(e.children[k]));
      return out;
    }
    case SExpr::kSpinOp: {
      if (particle != kSpin) throw ExprError{"spin operator in a fermionic expression"};
      int const i = site(e.site);
      if (i < 0 || i >= number_sites) throw ExprError{"site index " + std::to_string(i) + " out of range"};
      double const f = e.prefix == 'S' ? 0.5 : 1.0;
      Poly p;
      switch (e.op) {
      case 'x': p = poly_add(spin_plus(i), spin_minus(i)); break;
      case 'y': p = poly_add(poly_scale(spin_plus(i), cplx(0, -1)), poly_scale(spin_minus(i), cplx(0, 1))); break;
      case 'z': p = spin_z(i); break;
      case '+': p = spin_plus(i); break;
      default: p = spin_minus(i); break;
      }
      return poly_scale(p, f);
    }
    case SExpr::kFermionOp: {
      if (particle == kSpin) throw ExprError{"fermionic operator in a spin expression"};
      int const i = site(e.site);
      if (i < 0 || i >= number_sites) throw ExprError{"site index " + std::to_string(i) + " out of range"};
      int bit = i;
      if (particle == kSpinfulFermion) {
        if (e.spin < 0) throw ExprError{"spinful fermion operator without a spin label"};
        bit = i + (e.spin == 1 ? number_sites : 0);
      } else if (e.spin >= 0) {
        throw ExprError{"spin label on a spinless fermion operator"};
      }
      if (e.op == 'c') return fermion_create(bit);
      if (e.op == 'a') return fermion_annihilate(bit);
      return fermion_number(bit);
    }
    }
    return {};
  }
};

long to_int(NodeP const &n, char const *what) {
  if (!n || n->kind != Node::kScalar) throw YamlError{std::string("expected an integer for ") + what};
  char *end = nullptr;
  long const v = std::strtol(n->text.c_str(), &end, 10);
  if (end == n->text.c_str() || *end != '\0') throw YamlError{std::string("expected an integer for ") + what};
  return v;
}

} // namespace

struct ModelSymmetry {
  std::vector<int32_t> permutation;
  int32_t sector = 0;
};

struct ModelOperator {
  std::string name;
  Poly terms;
};

} // namespace lsamd

struct ls_amd_model {
  int particle = lsamd::kSpin;
  int number_sites = 0;
  int hamming_weight = -1; // spins
  int spin_inversion = 0;
  int number_up = -1, number_down = -1; // fermions (number_down < 0: spinless or total only)
  int number_particles = -1;            // spinful with a total count only
  std::vector<lsamd::ModelSymmetry> symmetries;
  bool has_hamiltonian = false;
  lsamd::ModelOperator hamiltonian;
  std::vector<lsamd::ModelOperator> observables;
};

namespace lsamd {
namespace {

int particle_from_text(std::string const &t) {
  if (t == "spin" || t == "spin-1/2") return kSpin;
  if (t == "spinless" || t == "spinless-fermion" || t == "spinless fermion") return kSpinlessFermion;
  if (t == "spinful" || t == "spinful-fermion" || t == "spinful fermion") return kSpinfulFermion;
  throw YamlError{"invalid particle type: " + t};
}

// Operator: {name?, terms: [{expression, sites?, particle?}, ...]} (Yaml.hs:38-41).
ModelOperator read_operator(ls_amd_model const &m, NodeP const &node) {
  if (!node || node->kind != Node::kMap) throw YamlError{"an operator must be a mapping"};
  ModelOperator op;
  if (NodeP nm = node->get("name"); nm && nm->kind == Node::kScalar) op.name = nm->text;
  NodeP terms = node->get("terms");
  if (!terms || terms->kind != Node::kSeq || terms->items.empty())
    throw YamlError{"an operator needs a non-empty 'terms' list"};
  for (NodeP const &t : terms->items) {
    if (!t || t->kind != Node::kMap) throw YamlError{"each term must be a mapping"};
    NodeP ex = t->get("expression");
    if (!ex || ex->kind != Node::kScalar) throw YamlError{"a term needs an 'expression'"};
    ExprParser parser;
    SExpr const e = parser.parse(ex->text);
    int p = particle_of(e);
    if (NodeP pt = t->get("particle"); pt && pt->kind == Node::kScalar) {
      int const want = particle_from_text(pt->text);
      if (p >= 0 && p != want) throw ExprError{"expression does not match its particle type"};
      p = want;
    }
    if (p >= 0 && p != m.particle) throw ExprError{"expression particle type differs from the basis"};
    Lowering low{m.particle, m.number_sites, {}};
    NodeP sites = t->get("sites");
    if (!sites || sites->kind == Node::kNull) {
      op.terms = poly_add(std::move(op.terms), low.lower(e));
      continue;
    }
    if (sites->kind != Node::kSeq) throw YamlError{"'sites' must be a list of site lists"};
    std::set<int> old;
    collect_sites(e, old);
    std::vector<int> const olds(old.begin(), old.end());
    for (NodeP const &tuple : sites->items) {
      if (!tuple || tuple->kind != Node::kSeq) throw YamlError{"'sites' must be a list of site lists"};
      if (tuple->items.size() != olds.size())
        throw ExprError{"wrong number of site indices: " + std::to_string(tuple->items.size()) + "; expected " +
                        std::to_string(olds.size())};
      low.remap.clear();
      for (size_t k = 0; k < olds.size(); ++k) low.remap[olds[k]] = static_cast<int>(to_int(tuple->items[k], "a site"));
      op.terms = poly_add(std::move(op.terms), low.lower(e));
    }
  }
  return op;
}

void read_basis(ls_amd_model &m, NodeP const &b) {
  if (!b || b->kind != Node::kMap) throw YamlError{"missing 'basis' mapping"};
  int particle = kSpin;
  if (NodeP p = b->get("particle"); p && p->kind == Node::kScalar) particle = particle_from_text(p->text);
  m.particle = particle;
  if (particle == kSpin) {
    m.number_sites = static_cast<int>(to_int(b->get("number_spins"), "number_spins"));
    if (m.number_sites <= 0 || m.number_sites > 64) throw YamlError{"invalid number of spins"};
    if (NodeP h = b->get("hamming_weight"); h && h->kind != Node::kNull)
      m.hamming_weight = static_cast<int>(to_int(h, "hamming_weight"));
    if (NodeP i = b->get("spin_inversion"); i && i->kind != Node::kNull)
      m.spin_inversion = static_cast<int>(to_int(i, "spin_inversion"));
    // checks of Basis.hs:328-345
    if (m.hamming_weight != -1 && (m.hamming_weight < 0 || m.hamming_weight > m.number_sites))
      throw YamlError{"invalid Hamming weight"};
    if (m.spin_inversion != 0 && m.spin_inversion != 1 && m.spin_inversion != -1)
      throw YamlError{"invalid spin inversion"};
    if (m.spin_inversion != 0 && m.hamming_weight == -1) m.hamming_weight = m.number_sites / 2;
    if (m.spin_inversion != 0 && 2 * m.hamming_weight != m.number_sites)
      throw YamlError{"spin inversion needs Hamming weight n / 2"};
  } else {
    m.number_sites = static_cast<int>(to_int(b->get("number_sites"), "number_sites"));
    int const bits = particle == kSpinfulFermion ? 2 * m.number_sites : m.number_sites;
    if (m.number_sites <= 0 || bits > 64) throw YamlError{"invalid number of sites"};
    NodeP np = b->get("number_particles");
    if (np && np->kind == Node::kScalar) {
      if (particle == kSpinfulFermion) m.number_particles = static_cast<int>(to_int(np, "number_particles"));
      else m.number_up = static_cast<int>(to_int(np, "number_particles"));
    } else if (np && np->kind == Node::kSeq) {
      if (particle != kSpinfulFermion || np->items.size() != 2)
        throw YamlError{"number_particles as [up, down] needs a spinful basis"};
      m.number_up = static_cast<int>(to_int(np->items[0], "number_particles"));
      m.number_down = static_cast<int>(to_int(np->items[1], "number_particles"));
    }
  }
  if (NodeP sy = b->get("symmetries"); sy && sy->kind != Node::kNull) {
    if (sy->kind != Node::kSeq) throw YamlError{"'symmetries' must be a list"};
    for (NodeP const &s : sy->items) {
      if (!s || s->kind != Node::kMap) throw YamlError{"each symmetry must be a mapping"};
      NodeP perm = s->get("permutation");
      if (!perm || perm->kind != Node::kSeq) throw YamlError{"a symmetry needs a 'permutation' list"};
      ModelSymmetry ms;
      for (NodeP const &v : perm->items) ms.permutation.push_back(static_cast<int32_t>(to_int(v, "a permutation")));
      if (static_cast<int>(ms.permutation.size()) != m.number_sites)
        throw YamlError{"permutation length differs from the number of sites"};
      ms.sector = static_cast<int32_t>(to_int(s->get("sector"), "sector"));
      m.symmetries.push_back(std::move(ms));
    }
  }
}

} // namespace
} // namespace lsamd

using namespace lsamd;

extern "C" ls_amd_model *ls_amd_model_from_yaml(char const *text) {
  if (text == nullptr) {
    fail("ls_amd_model_from_yaml: null text");
    return nullptr;
  }
  try {
    Yaml y;
    NodeP root = y.parse(text);
    if (!root || root->kind != Node::kMap) throw YamlError{"the model file must be a mapping"};
    auto m = std::make_unique<ls_amd_model>();
    read_basis(*m, root->get("basis"));
    if (NodeP h = root->get("hamiltonian"); h && h->kind != Node::kNull) {
      m->hamiltonian = read_operator(*m, h);
      m->has_hamiltonian = true;
    }
    if (NodeP obs = root->get("observables"); obs && obs->kind != Node::kNull) {
      if (obs->kind != Node::kSeq) throw YamlError{"'observables' must be a list"};
      for (NodeP const &o : obs->items) m->observables.push_back(read_operator(*m, o));
    }
    return m.release();
  } catch (YamlError const &e) {
    fail("model file: " + e.msg);
  } catch (ExprError const &e) {
    fail("expression: " + e.msg);
  }
  return nullptr;
}

extern "C" void ls_amd_model_destroy(ls_amd_model *m) { delete m; }

extern "C" int ls_amd_model_basis_info(ls_amd_model const *m, int *info) {
  if (m == nullptr || info == nullptr) return fail("null argument") ? 0 : -1;
  info[0] = m->particle;
  info[1] = m->number_sites;
  info[2] = m->hamming_weight;
  info[3] = m->spin_inversion;
  info[4] = m->number_up;
  info[5] = m->number_down;
  info[6] = m->number_particles;
  info[7] = static_cast<int>(m->symmetries.size());
  info[8] = m->has_hamiltonian ? 1 : 0;
  info[9] = static_cast<int>(m->observables.size());
  return 0;
}

extern "C" int ls_amd_model_symmetry(ls_amd_model const *m, int index, int32_t *permutation, int32_t *sector) {
  if (m == nullptr || permutation == nullptr || sector == nullptr) return fail("null argument") ? 0 : -1;
  if (index < 0 || index >= static_cast<int>(m->symmetries.size())) return fail("symmetry index out of range") ? 0 : -1;
  std::copy(m->symmetries[index].permutation.begin(), m->symmetries[index].permutation.end(), permutation);
  *sector = m->symmetries[index].sector;
  return 0;
}

namespace {
ModelOperator const *pick_operator(ls_amd_model const *m, int which) {
  if (which < 0) return m->has_hamiltonian ? &m->hamiltonian : nullptr;
  if (which < static_cast<int>(m->observables.size())) return &m->observables[which];
  return nullptr;
}
} // namespace

extern "C" int64_t ls_amd_model_number_terms(ls_amd_model const *m, int which, int diagonal) {
  if (m == nullptr) return fail("null argument") ? 0 : -1;
  ModelOperator const *op = pick_operator(m, which);
  if (op == nullptr) return fail("no such operator in the model") ? 0 : -1;
  int64_t k = 0;
  for (auto const &[key, v] : op->terms)
    if ((key.x == 0) == (diagonal != 0)) ++k;
  return k;
}

extern "C" int ls_amd_model_terms(ls_amd_model const *m, int which, int diagonal, double *v, uint64_t *mm,
                                  uint64_t *l, uint64_t *r, uint64_t *x, uint64_t *s) {
  if (m == nullptr || v == nullptr || mm == nullptr || l == nullptr || r == nullptr || x == nullptr || s == nullptr)
    return fail("null argument") ? 0 : -1;
  ModelOperator const *op = pick_operator(m, which);
  if (op == nullptr) return fail("no such operator in the model") ? 0 : -1;
  int64_t k = 0;
  for (auto const &[key, val] : op->terms) {
    if ((key.x == 0) != (diagonal != 0)) continue;
    v[2 * k] = val.real();
    v[2 * k + 1] = val.imag();
    mm[k] = key.m;
    l[k] = key.l;
    r[k] = key.r;
    x[k] = key.x;
    s[k] = key.s;
    ++k;
  }
  return 0;
}

namespace {
ls_hs_permutation_group *model_group(ls_amd_model const *m) {
  if (m->symmetries.empty()) return nullptr;
  std::vector<int32_t> perms, sectors;
  for (auto const &sy : m->symmetries) {
    perms.insert(perms.end(), sy.permutation.begin(), sy.permutation.end());
    sectors.push_back(sy.sector);
  }
  return ls_amd_permutation_group_create(m->number_sites, static_cast<int>(m->symmetries.size()), perms.data(),
                                         sectors.data());
}
} // namespace

extern "C" ls_amd_basis *ls_amd_model_create_basis(ls_amd_model const *m) {
  if (m == nullptr) {
    fail("null argument");
    return nullptr;
  }
  ls_hs_permutation_group *g = nullptr;
  if (!m->symmetries.empty()) {
    g = model_group(m);
    if (g == nullptr) return nullptr;
  }
  ls_amd_basis *b = nullptr;
  if (m->particle == kSpin) {
    b = ls_amd_basis_create(m->number_sites, m->hamming_weight, m->spin_inversion, g);
  } else if (m->particle == kSpinlessFermion) {
    b = ls_amd_fermion_basis_create(m->number_sites, m->number_up, -1, g);
  } else {
    if (m->number_particles >= 0 || m->number_up < 0) {
      fail("spinful fermion bases need number_particles as [up, down] here");
    } else {
      b = ls_amd_fermion_basis_create(m->number_sites, m->number_up, m->number_down, g);
    }
  }
  if (g != nullptr) ls_amd_permutation_group_destroy(g);
  return b;
}

extern "C" ls_amd_operator *ls_amd_model_create_operator(ls_amd_model const *m, int which, ls_amd_basis const *b) {
  if (m == nullptr || b == nullptr) {
    fail("null argument");
    return nullptr;
  }
  ModelOperator const *op = pick_operator(m, which);
  if (op == nullptr) {
    fail("no such operator in the model");
    return nullptr;
  }
  int const bits = m->particle == kSpinfulFermion ? 2 * m->number_sites : m->number_sites;
  struct Arrays {
    std::vector<ls_hs_scalar> v;
    std::vector<uint64_t> mm, l, r, x, s;
    ls_hs_nonbranching_terms t{};
  } d, o;
  for (auto const &[key, val] : op->terms) {
    Arrays &a = key.x == 0 ? d : o;
    ls_hs_scalar c;
    c._real = val.real();
    c._imag = val.imag();
    a.v.push_back(c);
    a.mm.push_back(key.m);
    a.l.push_back(key.l);
    a.r.push_back(key.r);
    a.x.push_back(key.x);
    a.s.push_back(key.s);
  }
  for (Arrays *a : {&d, &o}) {
    a->t.number_terms = static_cast<int>(a->v.size());
    a->t.number_bits = bits;
    a->t.v = a->v.data();
    a->t.m = a->mm.data();
    a->t.l = a->l.data();
    a->t.r = a->r.data();
    a->t.x = a->x.data();
    a->t.s = a->s.data();
  }
  return ls_amd_operator_create(b, d.t.number_terms > 0 ? &d.t : nullptr, o.t.number_terms > 0 ? &o.t : nullptr);
}
