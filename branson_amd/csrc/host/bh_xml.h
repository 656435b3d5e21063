// bh_xml.h -- a small XML element reader for Branson input decks.
//
// Branson decks (src/input.h:105-449) are plain element trees: nested
// elements whose leaves carry text.  This reader handles elements, text,
// comments, processing instructions, attributes (skipped) and self-closing
// tags -- the subset those decks use -- and reports malformed input with a
// line number instead of guessing.
#pragma once

#include <memory>
#include <string>
#include <vector>

namespace bhx {

struct Node {
  std::string name;
  std::string text;  // concatenated character data, entities decoded
  std::vector<std::unique_ptr<Node>> children;

  const Node* child(const std::string& n) const;            // first child named n
  std::vector<const Node*> children_named(const std::string& n) const;
  std::string child_text(const std::string& n) const;       // trimmed, "" if absent
};

// Parses `src`; on failure returns nullptr and sets *err.
std::unique_ptr<Node> parse(const std::string& src, std::string* err);
std::unique_ptr<Node> parse_file(const std::string& path, std::string* err);

std::string trim(const std::string& s);

}  // namespace bhx
