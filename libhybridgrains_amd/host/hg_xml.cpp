// hg_xml.cpp -- recursive-descent reader for the scene XML subset.
#include "hg_xml.h"

#include <cctype>
#include <cstring>
#include <fstream>
#include <sstream>

namespace hg {
namespace xml {

const std::string *Node::attr(const std::string &key) const {
  for (const auto &a : attrs)
    if (a.first == key) return &a.second;
  return nullptr;
}

const Node *Node::child(const std::string &key) const {
  for (const auto &c : children)
    if (c->name == key) return c.get();
  return nullptr;
}

std::vector<const Node *> Node::children_named(const std::string &key) const {
  std::vector<const Node *> out;
  for (const auto &c : children)
    if (c->name == key) out.push_back(c.get());
  return out;
}

namespace {

struct Reader {
  const std::string &s;
  size_t i = 0;
  std::string err;

  explicit Reader(const std::string &t) : s(t) {}

  bool eof() const { return i >= s.size(); }
  bool starts(const char *p) const { return s.compare(i, std::strlen(p), p) == 0; }
  void skip_ws() {
    while (!eof() && std::isspace(static_cast<unsigned char>(s[i]))) ++i;
  }
  bool fail(const std::string &m) {
    if (err.empty()) {
      size_t line = 1;
      for (size_t k = 0; k < i && k < s.size(); ++k) line += s[k] == '\n';
      err = m + " (line " + std::to_string(line) + ")";
    }
    return false;
  }
  static bool name_char(char c) {
    return std::isalnum(static_cast<unsigned char>(c)) || c == '_' || c == '-' || c == ':' ||
           c == '.';
  }
  bool read_name(std::string &out) {
    size_t b = i;
    while (!eof() && name_char(s[i])) ++i;
    if (i == b) return fail("expected a name");
    out.assign(s, b, i - b);
    return true;
  }
  // Skips comments, processing instructions and doctype declarations.
  bool skip_misc() {
    for (;;) {
      skip_ws();
      if (starts("<!--")) {
        size_t e = s.find("-->", i + 4);
        if (e == std::string::npos) return fail("unterminated comment");
        i = e + 3;
      } else if (starts("<?")) {
        size_t e = s.find("?>", i + 2);
        if (e == std::string::npos) return fail("unterminated declaration");
        i = e + 2;
      } else if (starts("<!")) {
        size_t e = s.find('>', i + 2);
        if (e == std::string::npos) return fail("unterminated declaration");
        i = e + 1;
      } else {
        return true;
      }
    }
  }
  static std::string unescape(const std::string &v) {
    std::string o;
    o.reserve(v.size());
    for (size_t k = 0; k < v.size(); ++k) {
      if (v[k] == '&') {
        static const struct { const char *ent; char c; } tab[] = {
            {"&lt;", '<'}, {"&gt;", '>'}, {"&amp;", '&'}, {"&quot;", '"'}, {"&apos;", '\''}};
        bool hit = false;
        for (const auto &t : tab) {
          const size_t n = std::strlen(t.ent);
          if (v.compare(k, n, t.ent) == 0) {
            o.push_back(t.c);
            k += n - 1;
            hit = true;
            break;
          }
        }
        if (!hit) o.push_back('&');
      } else {
        o.push_back(v[k]);
      }
    }
    return o;
  }
  bool element(Node &n) {
    if (!starts("<")) return fail("expected '<'");
    ++i;
    if (!read_name(n.name)) return false;
    for (;;) {
      skip_ws();
      if (eof()) return fail("unterminated start tag <" + n.name + ">");
      if (starts("/>")) {
        i += 2;
        return true;
      }
      if (s[i] == '>') {
        ++i;
        break;
      }
      std::string key;
      if (!read_name(key)) return false;
      skip_ws();
      if (eof() || s[i] != '=') return fail("expected '=' after attribute " + key);
      ++i;
      skip_ws();
      if (eof() || (s[i] != '"' && s[i] != '\'')) return fail("expected a quoted value");
      const char q = s[i++];
      size_t e = s.find(q, i);
      if (e == std::string::npos) return fail("unterminated attribute value");
      n.attrs.emplace_back(key, unescape(s.substr(i, e - i)));
      i = e + 1;
    }
    // content: child elements, comments and (ignored) text
    for (;;) {
      if (eof()) return fail("missing </" + n.name + ">");
      if (starts("<!--") || starts("<?") || starts("<!")) {
        if (!skip_misc()) return false;
        continue;
      }
      if (starts("</")) {
        i += 2;
        std::string close;
        if (!read_name(close)) return false;
        if (close != n.name) return fail("</" + close + "> closes <" + n.name + ">");
        skip_ws();
        if (eof() || s[i] != '>') return fail("expected '>'");
        ++i;
        return true;
      }
      if (s[i] == '<') {
        auto c = std::make_unique<Node>();
        if (!element(*c)) return false;
        n.children.push_back(std::move(c));
        continue;
      }
      ++i;  // text
    }
  }
};

}  // namespace

std::unique_ptr<Node> parse(const std::string &text, std::string *error) {
  Reader r(text);
  auto doc = std::make_unique<Node>();
  doc->name = "#document";
  for (;;) {
    if (!r.skip_misc()) break;
    if (r.eof()) break;
    auto c = std::make_unique<Node>();
    if (!r.element(*c)) break;
    doc->children.push_back(std::move(c));
  }
  if (!r.err.empty()) {
    if (error) *error = r.err;
    return nullptr;
  }
  return doc;
}

std::unique_ptr<Node> parse_file(const std::string &path, std::string *error) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    if (error) *error = "cannot open " + path;
    return nullptr;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str(), error);
}

}  // namespace xml
}  // namespace hg
