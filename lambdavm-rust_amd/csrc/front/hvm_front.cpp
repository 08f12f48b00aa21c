// hvm_front.cpp — see hvm_front.hpp for what each part restates.
#include "hvm_front.hpp"

#include <charconv>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <set>

namespace hvm {

// ---------------------------------------------------------------------------
// Tree helpers
// ---------------------------------------------------------------------------

std::unique_ptr<Tree> Tree::var(std::string n) {
  auto t = std::make_unique<Tree>(); t->kind = TreeKind::Var; t->nam = std::move(n); return t;
}
std::unique_ptr<Tree> Tree::ref(std::string n) {
  auto t = std::make_unique<Tree>(); t->kind = TreeKind::Ref; t->nam = std::move(n); return t;
}
std::unique_ptr<Tree> Tree::era() {
  auto t = std::make_unique<Tree>(); t->kind = TreeKind::Era; return t;
}
std::unique_ptr<Tree> Tree::numb(uint32_t v) {
  auto t = std::make_unique<Tree>(); t->kind = TreeKind::Num; t->num = v; return t;
}
std::unique_ptr<Tree> Tree::node(TreeKind k, std::unique_ptr<Tree> a, std::unique_ptr<Tree> b) {
  auto t = std::make_unique<Tree>(); t->kind = k; t->fst = std::move(a); t->snd = std::move(b); return t;
}

// ---------------------------------------------------------------------------
// Parser.  TSPL 0.0.13 is not vendored; its primitives are restated from its
// call sites in src/ast.rs:42-230: `skip_trivia` skips ASCII whitespace and
// `//` line comments, `parse_name` takes [A-Za-z0-9_.-/]+, `parse_u64`
// accepts an optional 0x/0b/0o radix prefix and `_` separators.
// ---------------------------------------------------------------------------

static bool is_name_char(int c) {
  return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') ||
         c == '_' || c == '.' || c == '-' || c == '/';
}

static bool is_alnum(int c) {
  return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9');
}

void Parser::expected(const std::string& what) {
  size_t lo = i_ > 20 ? i_ - 20 : 0;
  std::string ctx = in_.substr(lo, 40);
  throw ParseError("Expected " + what + " at index " + std::to_string(i_) + " near `" + ctx + "`");
}

void Parser::skip_trivia() {
  while (i_ < in_.size()) {
    char c = in_[i_];
    if (c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f') { ++i_; continue; }
    if (c == '/' && i_ + 1 < in_.size() && in_[i_ + 1] == '/') {
      while (i_ < in_.size() && in_[i_] != '\n') ++i_;
      if (i_ < in_.size()) ++i_;
      continue;
    }
    break;
  }
}

bool Parser::is_eof() {
  skip_trivia();
  return i_ >= in_.size();
}

void Parser::consume(const char* text) {
  skip_trivia();
  size_t n = strlen(text);
  if (in_.compare(i_, n, text) == 0 && i_ + n <= in_.size()) {
    i_ += n;
    return;
  }
  expected(std::string("'") + text + "'");
}

bool Parser::try_consume(const char* text) {
  size_t n = strlen(text);
  if (i_ + n <= in_.size() && in_.compare(i_, n, text) == 0) {
    i_ += n;
    return true;
  }
  return false;
}

std::string Parser::parse_name() {
  skip_trivia();
  size_t s = i_;
  while (i_ < in_.size() && is_name_char((unsigned char)in_[i_])) ++i_;
  if (i_ == s) expected("name");
  return in_.substr(s, i_ - s);
}

// TSPL parse_u64 over a literal's text: radix prefix, digits and `_`, stops at
// the first other character (the rest of the literal is ignored).
static uint64_t tspl_parse_u64(const std::string& s, size_t pos, bool& ok) {
  int radix = 10;
  if (s.compare(pos, 2, "0x") == 0) { radix = 16; pos += 2; }
  else if (s.compare(pos, 2, "0b") == 0) { radix = 2; pos += 2; }
  else if (s.compare(pos, 2, "0o") == 0) { radix = 8; pos += 2; }
  uint64_t v = 0;
  size_t digits = 0;
  for (; pos < s.size(); ++pos) {
    int c = (unsigned char)s[pos], d;
    if (c == '_') continue;
    if (c >= '0' && c <= '9') d = c - '0';
    else if (c >= 'a' && c <= 'z') d = c - 'a' + 10;
    else if (c >= 'A' && c <= 'Z') d = c - 'A' + 10;
    else break;
    if (d >= radix) break;
    uint64_t nv = v * (uint64_t)radix + (uint64_t)d;
    if (nv / (uint64_t)radix != v && v != 0) { ok = false; return 0; }  // u64 overflow is an error
    v = nv;
    ++digits;
  }
  ok = digits > 0;
  return v;
}

uint32_t Parser::parse_numb_lit() {
  size_t s = i_;
  while (i_ < in_.size()) {
    int c = (unsigned char)in_[i_];
    if (is_alnum(c) || c == '+' || c == '-' || c == '.') ++i_; else break;
  }
  std::string num = in_.substr(s, i_ - s);
  if (num.find('.') != std::string::npos || num.find("inf") != std::string::npos ||
      num.find("NaN") != std::string::npos) {
    // Rust's f32 parser: correctly rounded decimal; accepts inf/NaN spellings.
    std::string t = num;
    const char* p = t.c_str();
    bool neg = false;
    if (*p == '+' || *p == '-') { neg = (*p == '-'); ++p; }
    float v;
    if (strcmp(p, "inf") == 0 || strcmp(p, "infinity") == 0 || strcmp(p, "Infinity") == 0) {
      v = neg ? -INFINITY : INFINITY;
    } else if (strcmp(p, "NaN") == 0 || strcmp(p, "nan") == 0) {
      v = NAN;
    } else {
      char* end = nullptr;
      v = strtof(t.c_str(), &end);
      if (end == t.c_str() || *end != '\0') {
        i_ = s;
        expected("number literal");
      }
    }
    return hvm_new_f24_literal(v);
  }
  bool ok = false;
  if (!num.empty() && (num[0] == '+' || num[0] == '-')) {
    uint64_t u = tspl_parse_u64(num, 1, ok);
    if (!ok) { i_ = s; expected("numeric digit"); }
    int32_t val = (int32_t)(uint32_t)u;
    return hvm_new_i24(num[0] == '-' ? (int32_t)(0u - (uint32_t)val) : val);
  }
  uint64_t u = tspl_parse_u64(num, 0, ok);
  if (!ok) { i_ = s; expected("numeric digit"); }
  return hvm_new_u24((uint32_t)u);
}

uint32_t Parser::parse_numb_sym() {
  consume("[");
  if (try_consume("u24")) { consume("]"); return hvm_new_sym(HVM_TY_U24); }
  if (try_consume("i24")) { consume("]"); return hvm_new_sym(HVM_TY_I24); }
  if (try_consume("f24")) { consume("]"); return hvm_new_sym(HVM_TY_F24); }
  // Order matters: the reference tries these prefixes in this order.
  static const struct { const char* s; uint32_t op; } ops[] = {
    {"+", HVM_OP_ADD}, {"-", HVM_OP_SUB}, {":-", HVM_FP_SUB}, {"*", HVM_OP_MUL},
    {"/", HVM_OP_DIV}, {":/", HVM_FP_DIV}, {"%", HVM_OP_REM}, {":%", HVM_FP_REM},
    {"=", HVM_OP_EQ}, {"!", HVM_OP_NEQ}, {"<<", HVM_OP_SHL}, {":<<", HVM_FP_SHL},
    {">>", HVM_OP_SHR}, {":>>", HVM_FP_SHR}, {"<", HVM_OP_LT}, {">", HVM_OP_GT},
    {"&", HVM_OP_AND}, {"|", HVM_OP_OR}, {"^", HVM_OP_XOR},
  };
  uint32_t op = 0;
  bool found = false;
  for (auto& o : ops) {
    if (try_consume(o.s)) { op = hvm_new_sym(o.op); found = true; break; }
  }
  if (!found) expected("operator symbol");
  skip_trivia();
  uint32_t num = op;
  if (peek_one() != ']') num = hvm_partial(op, parse_numb_lit());
  consume("]");
  return num;
}

uint32_t Parser::parse_numb() {
  skip_trivia();
  if (peek_one() == '[') return parse_numb_sym();
  return parse_numb_lit();
}

std::unique_ptr<Tree> Parser::parse_tree() {
  skip_trivia();
  int c = peek_one();
  auto binary = [&](TreeKind k, const char* close) {
    auto fst = parse_tree();
    skip_trivia();
    auto snd = parse_tree();
    consume(close);
    return Tree::node(k, std::move(fst), std::move(snd));
  };
  switch (c) {
    case '(': ++i_; return binary(TreeKind::Con, ")");
    case '{': ++i_; return binary(TreeKind::Dup, "}");
    case '$': ++i_; consume("("); return binary(TreeKind::Opr, ")");
    case '?': ++i_; consume("("); return binary(TreeKind::Swi, ")");
    case '@': ++i_; return Tree::ref(parse_name());
    case '*': ++i_; return Tree::era();
    default: break;
  }
  if (c >= 0 && strchr("0123456789+-[", c) && c != 0) return Tree::numb(parse_numb());
  return Tree::var(parse_name());
}

Net Parser::parse_net() {
  Net net;
  net.root = parse_tree();
  skip_trivia();
  while (peek_one() == '&') {
    consume("&");
    bool par = false;
    if (peek_one() == '!') { consume("!"); par = true; }
    Redex r;
    r.par = par;
    r.fst = parse_tree();
    consume("~");
    r.snd = parse_tree();
    net.rbag.push_back(std::move(r));
    skip_trivia();
  }
  return net;
}

AstBook Parser::parse_book() {
  AstBook book;
  while (!is_eof()) {
    consume("@");
    std::string name = parse_name();
    consume("=");
    Net net = parse_net();
    book.defs[name] = std::move(net);  // a later def with the same name replaces it
  }
  return book;
}

AstBook parse_book(const std::string& code) {
  Parser p(code);
  return p.parse_book();
}

// ---------------------------------------------------------------------------
// Builder (src/ast.rs:432-559).  Nodes are numbered in pre-order, the root
// tree first and then each redex (fst, then snd); vars by first occurrence.
// ---------------------------------------------------------------------------

struct BuildCtx {
  Def* def;
  const std::map<std::string, uint32_t>* fids;
  std::map<std::string, uint32_t> vars;
};

static hvm_port build_tree(const Tree& t, BuildCtx& cx) {
  switch (t.kind) {
    case TreeKind::Var: {
      auto it = cx.vars.find(t.nam);
      if (it == cx.vars.end()) {
        uint32_t id = (uint32_t)cx.vars.size();
        cx.vars.emplace(t.nam, id);
        cx.def->vars += 1;
        return hvm_new_port(HVM_VAR, id);
      }
      return hvm_new_port(HVM_VAR, it->second);
    }
    case TreeKind::Ref: {
      auto it = cx.fids->find(t.nam);
      if (it == cx.fids->end()) throw std::runtime_error("Unbound definition: " + t.nam);
      return hvm_new_port(HVM_REF, it->second);
    }
    case TreeKind::Era: return hvm_new_port(HVM_ERA, 0);
    case TreeKind::Num: return hvm_new_port(HVM_NUM, t.num);
    default: break;
  }
  hvm_tag tag = t.kind == TreeKind::Con ? HVM_CON : t.kind == TreeKind::Dup ? HVM_DUP
              : t.kind == TreeKind::Opr ? HVM_OPR : HVM_SWI;
  if (t.kind == TreeKind::Dup) cx.def->safe = false;
  size_t index = cx.def->node.size();
  cx.def->node.push_back(0);
  hvm_port p1 = build_tree(*t.fst, cx);
  hvm_port p2 = build_tree(*t.snd, cx);
  cx.def->node[index] = hvm_new_pair(p1, p2);
  return hvm_new_port(tag, (uint32_t)index);
}

static void direct_dependencies(const Tree& t, std::set<std::string>& acc) {
  std::vector<const Tree*> stack{&t};
  while (!stack.empty()) {
    const Tree* c = stack.back();
    stack.pop_back();
    if (c->kind == TreeKind::Ref) acc.insert(c->nam);
    if (c->fst) { stack.push_back(c->fst.get()); stack.push_back(c->snd.get()); }
  }
}

Book build_book(const AstBook& ast) {
  std::map<std::string, uint32_t> name_to_fid;
  std::map<uint32_t, std::string> fid_to_name;
  name_to_fid["main"] = 0;
  fid_to_name[0] = "main";
  for (auto& kv : ast.defs) {
    if (kv.first != "main") {
      uint32_t fid = (uint32_t)name_to_fid.size();
      fid_to_name[fid] = kv.first;
      name_to_fid[kv.first] = fid;
    }
  }
  Book book;
  for (auto& kv : fid_to_name) {
    auto it = ast.defs.find(kv.second);
    if (it == ast.defs.end()) throw std::runtime_error("missing `@main` definition");
    Def def;
    def.name = kv.second;
    BuildCtx cx{&def, &name_to_fid, {}};
    const Net& net = it->second;
    def.root = build_tree(*net.root, cx);
    for (auto& rx : net.rbag) {
      size_t index = def.rbag.size();
      def.rbag.push_back(0);
      hvm_port p1 = build_tree(*rx.fst, cx);
      hvm_port p2 = build_tree(*rx.snd, cx);
      hvm_pair pr = hvm_new_pair(p1, p2);
      def.rbag[index] = rx.par ? hvm_set_par_flag(pr) : pr;
    }
    book.defs.push_back(std::move(def));
  }
  // Safety propagation (src/ast.rs:571-634): every def that (transitively)
  // references an unsafe def becomes unsafe.
  std::map<std::string, std::set<std::string>> dependents;
  for (auto& kv : ast.defs) dependents[kv.first];
  for (auto& kv : ast.defs) {
    std::set<std::string> deps;
    direct_dependencies(*kv.second.root, deps);
    for (auto& rx : kv.second.rbag) {
      direct_dependencies(*rx.fst, deps);
      direct_dependencies(*rx.snd, deps);
    }
    for (auto& d : deps) {
      auto dit = dependents.find(d);
      if (dit == dependents.end())
        throw std::runtime_error("global definition depends on undeclared reference");
      dit->second.insert(kv.first);
    }
  }
  std::vector<std::string> stack;
  for (auto& kv : ast.defs) {
    if (!book.defs[name_to_fid[kv.first]].safe) {
      for (auto& n : dependents[kv.first]) stack.push_back(n);
    }
  }
  while (!stack.empty()) {
    std::string cur = stack.back();
    stack.pop_back();
    Def& d = book.defs[name_to_fid[cur]];
    if (!d.safe) continue;
    d.safe = false;
    for (auto& n : dependents[cur]) stack.push_back(n);
  }
  return book;
}

// ---------------------------------------------------------------------------
// Wire format (src/hvm.rs:957-1003): native-endian u32 words.
// ---------------------------------------------------------------------------

std::vector<uint32_t> to_buffer(const Book& book) {
  std::vector<uint32_t> w;
  w.push_back((uint32_t)book.defs.size());
  for (size_t fid = 0; fid < book.defs.size(); ++fid) {
    const Def& d = book.defs[fid];
    w.push_back((uint32_t)fid);
    if (d.name.size() >= 256) throw std::runtime_error("Name too long: " + d.name);
    uint32_t name[64];
    memset(name, 0, sizeof(name));
    memcpy(name, d.name.data(), d.name.size());
    w.insert(w.end(), name, name + 64);
    w.push_back(d.safe ? 1u : 0u);
    w.push_back((uint32_t)d.rbag.size());
    w.push_back((uint32_t)d.node.size());
    w.push_back(d.vars);
    w.push_back(d.root);
    for (hvm_pair p : d.rbag) { w.push_back((uint32_t)p); w.push_back((uint32_t)(p >> 32)); }
    for (hvm_pair p : d.node) { w.push_back((uint32_t)p); w.push_back((uint32_t)(p >> 32)); }
  }
  return w;
}

bool load_buffer(const uint32_t* buf, size_t n_words, Book& out, std::string& err) {
  size_t i = 0;
  auto need = [&](size_t k) {
    if (n_words != (size_t)-1 && i + k > n_words) { err = "book buffer truncated"; return false; }
    return true;
  };
  if (!need(1)) return false;
  uint32_t defs_len = buf[i++];
  if (defs_len > HVM_MAX_DEFS) { err = "book has too many defs: " + std::to_string(defs_len); return false; }
  out.defs.clear();
  out.defs.resize(defs_len);
  std::vector<bool> seen(defs_len, false);
  for (uint32_t k = 0; k < defs_len; ++k) {
    if (!need(1 + 64 + 5)) return false;
    uint32_t fid = buf[i++];
    if (fid >= defs_len || seen[fid]) { err = "book has a bad fid: " + std::to_string(fid); return false; }
    seen[fid] = true;
    Def& d = out.defs[fid];
    char name[257];
    memcpy(name, buf + i, 256);
    name[256] = 0;
    d.name = name;
    i += 64;
    d.safe = buf[i++] != 0;
    uint32_t rlen = buf[i++], nlen = buf[i++];
    d.vars = buf[i++];
    if (rlen > HVM_DEF_RBAG_LEN) { err = "def '" + d.name + "' has too many redexes: " + std::to_string(rlen); return false; }
    if (nlen > HVM_DEF_NODE_LEN) { err = "def '" + d.name + "' has too many nodes: " + std::to_string(nlen); return false; }
    d.root = buf[i++];
    if (!need(2 * ((size_t)rlen + nlen))) return false;
    d.rbag.resize(rlen);
    d.node.resize(nlen);
    for (uint32_t r = 0; r < rlen; ++r, i += 2) d.rbag[r] = ((uint64_t)buf[i + 1] << 32) | buf[i];
    for (uint32_t n = 0; n < nlen; ++n, i += 2) d.node[n] = ((uint64_t)buf[i + 1] << 32) | buf[i];
  }
  return true;
}

// ---------------------------------------------------------------------------
// Printers
// ---------------------------------------------------------------------------

static const char* op_text(uint32_t op) {
  switch (op) {
    case HVM_OP_ADD: return "+";   case HVM_OP_SUB: return "-";   case HVM_FP_SUB: return ":-";
    case HVM_OP_MUL: return "*";   case HVM_OP_DIV: return "/";   case HVM_FP_DIV: return ":/";
    case HVM_OP_REM: return "%";   case HVM_FP_REM: return ":%";  case HVM_OP_EQ:  return "=";
    case HVM_OP_NEQ: return "!";   case HVM_OP_LT:  return "<";   case HVM_OP_GT:  return ">";
    case HVM_OP_AND: return "&";   case HVM_OP_OR:  return "|";   case HVM_OP_XOR: return "^";
    case HVM_OP_SHL: return "<<";  case HVM_FP_SHL: return ":<<"; case HVM_OP_SHR: return ">>";
    case HVM_FP_SHR: return ":>>"; default: return nullptr;
  }
}

static std::string sym_text(uint32_t sym) {
  switch (sym) {
    case HVM_TY_U24: return "[u24]";
    case HVM_TY_I24: return "[i24]";
    case HVM_TY_F24: return "[f24]";
    default: break;
  }
  const char* o = op_text(sym);
  return o ? std::string("[") + o + "]" : std::string("[?]");
}

static std::string partial_text(uint32_t numb) {
  char b[64];
  const char* o = op_text(hvm_get_typ(numb));
  snprintf(b, sizeof b, "[%s0x%07X]", o ? o : "?", hvm_get_u24(numb));
  return b;
}

// Rust's `{:?}` for f32: the shortest round-trip digits, in plain decimal
// notation (at least one fractional digit) when 1e-4 <= |x| < 1e16 or x == 0,
// else in `<d>[.<ddd>]e<exp>` notation.
static std::string rust_debug_f32(float v) {
  char sci[64];
  auto r = std::to_chars(sci, sci + sizeof sci, v, std::chars_format::scientific);
  *r.ptr = 0;
  // sci looks like "-1.2345e+07" or "5e-06"
  std::string s(sci);
  bool neg = false;
  if (!s.empty() && s[0] == '-') { neg = true; s = s.substr(1); }
  size_t e = s.find('e');
  std::string mant = s.substr(0, e);
  int exp = atoi(s.c_str() + e + 1);
  std::string digits;
  for (char c : mant) if (c != '.') digits.push_back(c);
  // strip trailing zeros of the digit string (keep at least one digit)
  while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
  float a = v < 0 ? -v : v;
  std::string out;
  if (v == 0.0f || (a >= 1e-4f && a < 1e16f)) {
    // decimal: value = 0.d1d2... * 10^(exp+1)
    int point = exp + 1;  // number of digits before the decimal point
    if (v == 0.0f) { digits = "0"; point = 1; }
    std::string ip, fp;
    if (point <= 0) {
      ip = "0";
      fp = std::string((size_t)(-point), '0') + digits;
    } else if ((size_t)point >= digits.size()) {
      ip = digits + std::string((size_t)point - digits.size(), '0');
      fp = "0";
    } else {
      ip = digits.substr(0, (size_t)point);
      fp = digits.substr((size_t)point);
    }
    out = ip + "." + fp;
  } else {
    out = digits.substr(0, 1);
    if (digits.size() > 1) out += "." + digits.substr(1);
    out += "e" + std::to_string(exp);
  }
  return (neg ? "-" : "") + out;
}

std::string show_numb(uint32_t numb) {
  switch (hvm_get_typ(numb)) {
    case HVM_TY_SYM: return sym_text(hvm_get_sym(numb));
    case HVM_TY_U24: return std::to_string(hvm_get_u24(numb));
    case HVM_TY_I24: {
      int32_t v = hvm_get_i24(numb);
      return (v >= 0 ? "+" : "") + std::to_string(v);
    }
    case HVM_TY_F24: {
      float f = hvm_get_f24(numb);
      if (std::isinf(f)) return f > 0 ? "+inf" : "-inf";
      if (std::isnan(f)) return "+NaN";
      return rust_debug_f32(f);
    }
    default: return partial_text(numb);
  }
}

std::string c_show_numb(uint32_t numb) {
  char b[64];
  switch (hvm_get_typ(numb)) {
    case HVM_TY_SYM: return sym_text(hvm_get_sym(numb));
    case HVM_TY_U24: snprintf(b, sizeof b, "%u", hvm_get_u24(numb)); return b;
    case HVM_TY_I24: snprintf(b, sizeof b, "%+d", hvm_get_i24(numb)); return b;
    case HVM_TY_F24: {
      float f = hvm_get_f24(numb);
      if (std::isinf(f)) return std::signbit(f) ? "-inf" : "+inf";
      if (std::isnan(f)) return "+NaN";
      snprintf(b, sizeof b, "%.7e", (double)f);
      return b;
    }
    default: return partial_text(numb);
  }
}

std::string Tree::show() const {
  // Iterative to survive deep right-nested lists.
  std::string out;
  std::vector<std::pair<const Tree*, const char*>> st;
  st.push_back({this, nullptr});
  while (!st.empty()) {
    auto [t, lit] = st.back();
    st.pop_back();
    if (lit) { out += lit; continue; }
    switch (t->kind) {
      case TreeKind::Var: out += t->nam; break;
      case TreeKind::Ref: out += "@" + t->nam; break;
      case TreeKind::Era: out += "*"; break;
      case TreeKind::Num: out += show_numb(t->num); break;
      default: {
        const char *open = "(", *close = ")";
        if (t->kind == TreeKind::Dup) { open = "{"; close = "}"; }
        else if (t->kind == TreeKind::Opr) { open = "$("; }
        else if (t->kind == TreeKind::Swi) { open = "?("; }
        out += open;
        st.push_back({nullptr, close});
        st.push_back({t->snd.get(), nullptr});
        st.push_back({nullptr, " "});
        st.push_back({t->fst.get(), nullptr});
      }
    }
  }
  return out;
}

void normalize_vars(Tree& root, std::map<std::string, size_t>& names) {
  std::vector<Tree*> st{&root};
  while (!st.empty()) {
    Tree* t = st.back();
    st.pop_back();
    if (t->kind == TreeKind::Var) {
      auto it = names.find(t->nam);
      size_t id;
      if (it == names.end()) { id = names.size(); names.emplace(t->nam, id); }
      else id = it->second;
      t->nam = "x" + std::to_string(id);
    } else if (t->fst) {
      st.push_back(t->snd.get());
      st.push_back(t->fst.get());
    }
  }
}

std::string canonical_tree_text(const std::string& printed) {
  Parser p(printed);
  auto t = p.parse_tree();
  std::map<std::string, size_t> names;
  normalize_vars(*t, names);
  return t->show();
}

}  // namespace hvm
