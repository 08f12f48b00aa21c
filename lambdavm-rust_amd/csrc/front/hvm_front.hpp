// hvm_front.hpp — host front end: `.hvm` text → Book → wire buffer, and back.
//
// This is the C++ restatement of the reference's Rust host layer for the GPU
// path (the reference cannot be built here: no Rust toolchain):
//   * parser          src/ast.rs:42-230   (TSPL 0.0.13 primitives restated)
//   * builder         src/ast.rs:432-559  (fid order, pre-order node numbering,
//                                          first-occurrence var numbering)
//   * safety          src/ast.rs:561-634  (DUP ⇒ unsafe, propagated to referrers)
//   * serializer      src/hvm.rs:957-1003 (Book::to_buffer wire format)
//   * loader          src/hvm.c:1408-1455 (book_load, incl. its size checks)
//   * printers        src/ast.rs:235-359  (Tree/Numb show, used for canonical NFs)
//                     src/hvm.c:1535-1699 (pretty_print_numb, the stdout format)
//   * canonical NF    tests/run.rs:112-146 (normalize_vars + Tree::show)
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

extern "C" {
#include "../hvm_core.h"
}

namespace hvm {

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// AST (src/ast.rs:10-35)
// ---------------------------------------------------------------------------

enum class TreeKind : uint8_t { Var, Ref, Era, Num, Con, Dup, Opr, Swi };

struct Tree {
  TreeKind kind = TreeKind::Era;
  std::string nam;        // Var / Ref
  uint32_t num = 0;       // Num (raw Numb)
  std::unique_ptr<Tree> fst, snd;

  static std::unique_ptr<Tree> var(std::string n);
  static std::unique_ptr<Tree> ref(std::string n);
  static std::unique_ptr<Tree> era();
  static std::unique_ptr<Tree> numb(uint32_t v);
  static std::unique_ptr<Tree> node(TreeKind k, std::unique_ptr<Tree> a, std::unique_ptr<Tree> b);

  std::string show() const;
};

struct Redex {
  bool par = false;
  std::unique_ptr<Tree> fst, snd;
};

struct Net {
  std::unique_ptr<Tree> root;
  std::vector<Redex> rbag;
};

struct AstBook {
  std::map<std::string, Net> defs;  // byte-wise ordered, like Rust's BTreeMap<String, _>
};

// ---------------------------------------------------------------------------
// Built book (src/hvm.rs:112-124)
// ---------------------------------------------------------------------------

struct Def {
  std::string name;
  bool safe = true;
  hvm_port root = 0;
  std::vector<hvm_pair> rbag;
  std::vector<hvm_pair> node;
  uint32_t vars = 0;
};

struct Book {
  std::vector<Def> defs;  // indexed by fid; fid 0 is @main
};

// Parser ------------------------------------------------------------------

class Parser {
 public:
  explicit Parser(const std::string& input) : in_(input), i_(0) {}
  AstBook parse_book();
  Net parse_net();
  std::unique_ptr<Tree> parse_tree();
  uint32_t parse_numb();
  void consume(const char* text);
  void skip_trivia();
  bool is_eof();
  size_t index() const { return i_; }

 private:
  uint32_t parse_numb_sym();
  uint32_t parse_numb_lit();
  std::string parse_name();
  bool try_consume(const char* text);
  int peek_one() const { return i_ < in_.size() ? (unsigned char)in_[i_] : -1; }
  [[noreturn]] void expected(const std::string& what);

  std::string in_;
  size_t i_;
};

AstBook parse_book(const std::string& code);

// Builder (src/ast.rs:532-559 + 561-634) ------------------------------------
// Throws std::runtime_error with the reference's panic text on a missing
// `@main` or an unbound reference.
Book build_book(const AstBook& ast);

// Wire format ---------------------------------------------------------------
std::vector<uint32_t> to_buffer(const Book& book);

// Loads a wire buffer.  Returns false and sets `err` like book_load does when
// a def exceeds the loader limits (src/hvm.c:1432-1440).
bool load_buffer(const uint32_t* buf, size_t n_words, Book& out, std::string& err);

// Compiled CALLs (hvm_spec.cpp) ---------------------------------------------
// Static interactions folded into one def's template: inlined CALLs and
// ANNIs, added to the counters at every expansion, and the canonical bytes of
// one expansion with them (the unspecialised def's own plus the static ones),
// which replace the specialised template's own count.
struct SpecCounts {
  uint32_t call = 0, anni = 0, bytes = 0;
};
// Every def's template reduced at load (REF~node CALLs inlined, same-tag node
// pairs annihilated, vars linked) within `node_budget` live nodes and
// `vars_max` vars; `counts[f]` receives def f's static counts.  Defs that do
// not reduce keep their template and zero counts.
Book specialize_book(const Book& book, uint32_t node_budget, uint32_t vars_max, std::vector<SpecCounts>& counts);

// Printers ------------------------------------------------------------------
std::string show_numb(uint32_t numb);      // Rust Numb::show (canonical form)
std::string c_show_numb(uint32_t numb);    // C pretty_print_numb (stdout contract)

// Canonical renaming of a `Result:` tree (tests/run.rs:135-146).
void normalize_vars(Tree& t, std::map<std::string, size_t>& names);

// Parses a printed tree (e.g. the text after "Result: " printed by the C
// runtime) and returns its canonical form.
std::string canonical_tree_text(const std::string& printed);

}  // namespace hvm
