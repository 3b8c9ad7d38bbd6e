// msp.hpp -- reader for the YAML subset hippoLBM's .msp decks use.
//
// Block maps and sequences by indentation, `- key: ...` sequence items that
// open a map, flow sequences / maps ([a, [b, c]], {k: v}), `|` block scalars
// (kept as text), comments.  Numbers may carry a unit suffix ("0.01 m"), as in
// regression-test/regression_lbm_lid_driven_cavity.msp; number() reads the
// leading value.  Anchors, tags and multi-document files are not supported.
#pragma once

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace msp {

struct Node {
  enum Kind { NUL, SCALAR, SEQ, MAP } kind = NUL;
  std::string text;                                 // SCALAR
  std::vector<Node> items;                          // SEQ
  std::vector<std::pair<std::string, Node>> fields;  // MAP, in file order

  bool is_null() const { return kind == NUL; }
  bool is_scalar() const { return kind == SCALAR; }
  bool is_seq() const { return kind == SEQ; }
  bool is_map() const { return kind == MAP; }
  // map lookup (nullptr when absent or not a map)
  const Node* find(const std::string& key) const;
  const Node& at(const std::string& key) const;  // throws when absent
  double number() const;                          // scalar -> double (unit suffix ignored)
  long integer() const;
  bool boolean() const;
  std::vector<double> numbers() const;            // flow/block sequence of numbers
};

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

Node parse(const std::string& text);
Node parse_file(const std::string& path);
// A deck with its `includes` (data/config/main-config.msp:11): each listed
// file, relative to the including file's directory, is read first, in order
// (recursively, at most 16 levels), and a top-level block of a later file
// replaces the same-named block of an earlier one; the including file's own
// blocks come last.  A missing include file is an error.
Node load_deck(const std::string& path);
// JSON rendering (scalars as strings), for tests and debugging
std::string to_json(const Node& n);

}  // namespace msp
