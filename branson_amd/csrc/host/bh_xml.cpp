// bh_xml.cpp -- see bh_xml.h.
#include "bh_xml.h"

#include <cctype>
#include <cstring>
#include <fstream>
#include <sstream>

namespace bhx {

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

const Node* Node::child(const std::string& n) const {
  for (const auto& c : children)
    if (c->name == n) return c.get();
  return nullptr;
}

std::vector<const Node*> Node::children_named(const std::string& n) const {
  std::vector<const Node*> out;
  for (const auto& c : children)
    if (c->name == n) out.push_back(c.get());
  return out;
}

std::string Node::child_text(const std::string& n) const {
  const Node* c = child(n);
  return c ? trim(c->text) : std::string();
}

namespace {

struct Parser {
  const std::string& s;
  size_t i = 0;
  std::string err;

  int line_at(size_t p) const {
    int l = 1;
    for (size_t k = 0; k < p && k < s.size(); ++k) l += (s[k] == '\n');
    return l;
  }
  bool fail(const std::string& m) {
    if (err.empty()) err = "line " + std::to_string(line_at(i)) + ": " + m;
    return false;
  }
  bool starts(const char* lit) const { return s.compare(i, std::strlen(lit), lit) == 0; }
  void skip_ws() {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
  }
  // skips <?...?>, <!--...-->, <!DOCTYPE ...>
  bool skip_misc() {
    while (true) {
      skip_ws();
      if (starts("<?")) {
        size_t e = s.find("?>", i);
        if (e == std::string::npos) return fail("unterminated processing instruction");
        i = e + 2;
      } else if (starts("<!--")) {
        size_t e = s.find("-->", i + 4);
        if (e == std::string::npos) return fail("unterminated comment");
        i = e + 3;
      } else if (starts("<!")) {
        size_t e = s.find('>', i);
        if (e == std::string::npos) return fail("unterminated declaration");
        i = e + 1;
      } else {
        return true;
      }
    }
  }
  static bool name_char(char c) {
    return std::isalnum((unsigned char)c) || c == '_' || c == '-' || c == '.' || c == ':';
  }
  std::string read_name() {
    size_t a = i;
    while (i < s.size() && name_char(s[i])) ++i;
    return s.substr(a, i - a);
  }
  static std::string decode(const std::string& t) {
    std::string o;
    o.reserve(t.size());
    for (size_t k = 0; k < t.size(); ++k) {
      if (t[k] != '&') {
        o += t[k];
        continue;
      }
      const size_t e = t.find(';', k);
      if (e == std::string::npos) {
        o += t[k];
        continue;
      }
      const std::string ent = t.substr(k + 1, e - k - 1);
      if (ent == "lt") o += '<';
      else if (ent == "gt") o += '>';
      else if (ent == "amp") o += '&';
      else if (ent == "quot") o += '"';
      else if (ent == "apos") o += '\'';
      else {
        o += t.substr(k, e - k + 1);
      }
      k = e;
    }
    return o;
  }

  std::unique_ptr<Node> element() {
    if (i >= s.size() || s[i] != '<') {
      fail("expected '<'");
      return nullptr;
    }
    ++i;
    auto n = std::make_unique<Node>();
    n->name = read_name();
    if (n->name.empty()) {
      fail("empty element name");
      return nullptr;
    }
    // attributes are not used by Branson decks: skip to the end of the tag
    while (i < s.size() && s[i] != '>' && !(s[i] == '/' && i + 1 < s.size() && s[i + 1] == '>')) {
      if (s[i] == '"' || s[i] == '\'') {
        const char q = s[i];
        const size_t e = s.find(q, i + 1);
        if (e == std::string::npos) {
          fail("unterminated attribute value");
          return nullptr;
        }
        i = e + 1;
      } else {
        ++i;
      }
    }
    if (i >= s.size()) {
      fail("unterminated tag <" + n->name + ">");
      return nullptr;
    }
    if (s[i] == '/') {  // self-closing
      i += 2;
      return n;
    }
    ++i;  // '>'
    std::string text;
    while (true) {
      if (i >= s.size()) {
        fail("missing </" + n->name + ">");
        return nullptr;
      }
      if (starts("<!--")) {
        const size_t e = s.find("-->", i + 4);
        if (e == std::string::npos) {
          fail("unterminated comment");
          return nullptr;
        }
        i = e + 3;
      } else if (starts("<![CDATA[")) {
        const size_t e = s.find("]]>", i + 9);
        if (e == std::string::npos) {
          fail("unterminated CDATA");
          return nullptr;
        }
        text += s.substr(i + 9, e - i - 9);
        i = e + 3;
      } else if (starts("</")) {
        i += 2;
        const std::string close = read_name();
        skip_ws();
        if (close != n->name || i >= s.size() || s[i] != '>') {
          fail("mismatched </" + close + "> for <" + n->name + ">");
          return nullptr;
        }
        ++i;
        n->text = decode(text);
        return n;
      } else if (s[i] == '<') {
        auto c = element();
        if (!c) return nullptr;
        n->children.push_back(std::move(c));
      } else {
        const size_t e = s.find('<', i);
        text += s.substr(i, (e == std::string::npos ? s.size() : e) - i);
        i = (e == std::string::npos) ? s.size() : e;
      }
    }
  }
};

}  // namespace

std::unique_ptr<Node> parse(const std::string& src, std::string* err) {
  Parser p{src};
  if (!p.skip_misc()) {
    if (err) *err = p.err;
    return nullptr;
  }
  auto root = p.element();
  if (root && !p.skip_misc()) root.reset();
  if (root) {
    p.skip_ws();
    if (p.i != src.size()) {
      p.fail("content after the root element");
      root.reset();
    }
  }
  if (!root && err) *err = p.err;
  return root;
}

std::unique_ptr<Node> parse_file(const std::string& path, std::string* err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    if (err) *err = "cannot open " + path;
    return nullptr;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  return parse(ss.str(), err);
}

}  // namespace bhx
