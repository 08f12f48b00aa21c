// hvm_spec.cpp — compiled CALLs: load-time specialisation of def templates.
//
// The reference compiles a def into a C function whose body performs some
// reductions of the expansion directly instead of creating the nodes and
// pushing the redexes (src/cmp.rs:36-104 compile_def, 106-275 the fast
// ANNI/SWITCH/OPER/COPY/VOID paths), adding the interactions it performed to
// ITRS (`tm->itrs += …`, src/cmp.rs:147,192,222,257,280).  On a GPU, per-def
// code diverges across the lanes of a wave (every lane may expand a different
// def), so this evaluator keeps one data-driven expansion loop and moves the
// reductions into the template instead: every def's template is reduced once,
// at load, as far as its own structure allows —
//
//   * CALL  `@f ~ (x y)` between a REF and a template node: the def's template
//     is inlined (its nodes and vars renumbered into the caller's, its redexes
//     added, its root linked to the node);
//   * ANNI  two template nodes of the same tag facing each other: both are
//     removed and their ports linked pairwise;
//   * LINK  a var facing anything: the var's other occurrence is replaced.
//
// Nothing else is reduced (no ERAS/COMM/OPER/SWIT, no CALL of a safe def
// against a DUP, which copies the REF): the reductions that remain depend on
// what the caller passes in.  Interaction nets are strongly confluent, so the
// net every expansion produces, and the number of interactions of each kind
// it takes to reach the normal form, are the same as without the
// specialisation; the evaluator adds the def's static CALL and ANNI counts
// (and their canonical bytes) to ITRS every time it expands the def, so ITRS
// and the per-kind histogram equal run-c's.
//
// Inlining stops at a node budget (live template nodes), 255 inlined CALLs
// and 4095 static ANNIs per def, so recursive defs terminate.  A def whose
// specialisation fails any check keeps its original template.
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "hvm_front.hpp"

namespace hvm {

namespace {

using u32 = uint32_t;
using u64 = uint64_t;

constexpr u32 kMaxCalls = 255;
constexpr u32 kMaxAnni = 4095;

struct Tmpl {
  std::vector<hvm_pair> node;
  std::vector<char> alive;
  u32 live = 0;
  hvm_port root = 0;
  std::vector<hvm_pair> work;   // redexes still to look at
  std::vector<hvm_pair> resid;  // redexes that stay in the template
  u32 vars = 0;
};

u64 canonical_call_bytes(const Def& d) {
  return 8ull * d.node.size() + 4ull * d.vars + 8ull * (d.rbag.size() + 1);
}

hvm_port shift(hvm_port p, u32 nb, u32 vb) {
  hvm_tag t = hvm_get_tag(p);
  if (t >= HVM_CON) return hvm_new_port(t, hvm_get_val(p) + nb);
  if (t == HVM_VAR) return hvm_new_port(t, hvm_get_val(p) + vb);
  return p;
}

// Connects the remaining occurrence of var `v` to `q`: a node slot or the root
// takes `q`; a redex side turns the redex into `q ~ other`, looked at again.
bool subst(Tmpl& t, u32 v, hvm_port q) {
  const hvm_port pv = hvm_new_port(HVM_VAR, v);
  if (t.root == pv) { t.root = q; return true; }
  for (size_t i = 0; i < t.node.size(); ++i) {
    if (!t.alive[i]) continue;
    hvm_pair& n = t.node[i];
    if (hvm_get_fst(n) == pv) { n = hvm_new_pair(q, hvm_get_snd(n)); return true; }
    if (hvm_get_snd(n) == pv) { n = hvm_new_pair(hvm_get_fst(n), q); return true; }
  }
  for (std::vector<hvm_pair>* bag : {&t.work, &t.resid}) {
    for (size_t i = 0; i < bag->size(); ++i) {
      const hvm_pair r = (*bag)[i];
      hvm_port other;
      if (hvm_get_fst(r) == pv) other = hvm_get_snd(r);
      else if (hvm_get_snd(r) == pv) other = hvm_get_fst(r);
      else continue;
      bag->erase(bag->begin() + (long)i);
      t.work.push_back(hvm_new_pair(q, other));
      return true;
    }
  }
  return false;
}

// Reduces the template; false if it turns out malformed (a var without a
// second occurrence, a var linked to itself).
bool reduce(Tmpl& t, const Book& book, u32 node_budget, SpecCounts& s) {
  for (u64 guard = 0; !t.work.empty(); ++guard) {
    if (guard > (1u << 22)) return false;
    const hvm_pair r = t.work.back();
    t.work.pop_back();
    hvm_port x = hvm_get_fst(r), y = hvm_get_snd(r);
    if (hvm_is_var(x) || hvm_is_var(y)) {
      if (x == y) return false;
      if (hvm_is_var(x) ? !subst(t, hvm_get_val(x), y) : !subst(t, hvm_get_val(y), x)) return false;
      continue;
    }
    if (hvm_get_tag(y) == HVM_REF) std::swap(x, y);
    if (hvm_get_tag(x) == HVM_REF && hvm_is_nod(y)) {
      const u32 fid = hvm_get_val(x) & 0xFFFFFFFu;
      if (fid >= book.defs.size()) return false;
      const Def& c = book.defs[fid];
      const bool copy = c.safe && hvm_get_tag(y) == HVM_DUP;
      if (!copy && t.live + c.node.size() <= node_budget && s.call < kMaxCalls &&
          s.anni + c.node.size() <= kMaxAnni) {
        const u32 nb = (u32)t.node.size(), vb = t.vars;
        for (hvm_pair p : c.node) {
          t.node.push_back(hvm_new_pair(shift(hvm_get_fst(p), nb, vb), shift(hvm_get_snd(p), nb, vb)));
          t.alive.push_back(1);
        }
        t.live += (u32)c.node.size();
        t.vars += c.vars;
        for (hvm_pair p : c.rbag)
          t.work.push_back(hvm_new_pair(shift(hvm_get_fst(p), nb, vb), shift(hvm_get_snd(p), nb, vb)));
        t.work.push_back(hvm_new_pair(shift(c.root, nb, vb), y));
        s.call += 1;
        s.bytes += (u32)canonical_call_bytes(c);
        continue;
      }
      t.resid.push_back(r);
      continue;
    }
    if (hvm_is_nod(x) && hvm_is_nod(y) && hvm_get_rule(x, y) == HVM_ANNI) {
      const u32 ia = hvm_get_val(x), ib = hvm_get_val(y);
      if (ia >= t.node.size() || ib >= t.node.size() || ia == ib || !t.alive[ia] || !t.alive[ib]) return false;
      const hvm_pair A = t.node[ia], B = t.node[ib];
      t.alive[ia] = t.alive[ib] = 0;
      t.live -= 2;
      t.work.push_back(hvm_new_pair(hvm_get_snd(A), hvm_get_snd(B)));
      t.work.push_back(hvm_new_pair(hvm_get_fst(A), hvm_get_fst(B)));
      s.anni += 1;
      s.bytes += 48;
      continue;
    }
    t.resid.push_back(r);
  }
  return true;
}

// Renumbers the live nodes (pre-order from the root, then from the redexes)
// and the vars (first occurrence), checking that every var occurs twice.
bool compact(const Tmpl& t, const Def& orig, u32 vars_max, Def& out) {
  std::vector<u32> nmap(t.node.size(), UINT32_MAX);
  std::vector<u32> order;
  std::vector<hvm_port> st;
  auto visit = [&](hvm_port root) {
    st.push_back(root);
    while (!st.empty()) {
      hvm_port p = st.back();
      st.pop_back();
      if (!hvm_is_nod(p)) continue;
      u32 i = hvm_get_val(p);
      if (i >= t.node.size() || !t.alive[i] || nmap[i] != UINT32_MAX) return false;
      nmap[i] = (u32)order.size();
      order.push_back(i);
      st.push_back(hvm_get_snd(t.node[i]));
      st.push_back(hvm_get_fst(t.node[i]));
    }
    return true;
  };
  if (!visit(t.root)) return false;
  for (hvm_pair r : t.resid)
    if (!visit(hvm_get_fst(r)) || !visit(hvm_get_snd(r))) return false;
  if (order.size() != t.live) return false;   // a live node nobody points to
  std::vector<u32> vmap(t.vars, UINT32_MAX);
  std::vector<u32> seen(t.vars, 0);
  u32 nv = 0;
  auto port = [&](hvm_port p, bool& ok) -> hvm_port {
    hvm_tag tg = hvm_get_tag(p);
    if (tg >= HVM_CON) return hvm_new_port(tg, nmap[hvm_get_val(p)]);
    if (tg == HVM_VAR) {
      u32 v = hvm_get_val(p);
      if (v >= t.vars) { ok = false; return p; }
      if (vmap[v] == UINT32_MAX) vmap[v] = nv++;
      seen[v]++;
      return hvm_new_port(HVM_VAR, vmap[v]);
    }
    return p;
  };
  bool ok = true;
  out.name = orig.name;
  out.safe = orig.safe;
  out.root = port(t.root, ok);
  out.node.clear();
  out.rbag.clear();
  for (u32 i : order) {
    hvm_port a = port(hvm_get_fst(t.node[i]), ok);
    hvm_port b = port(hvm_get_snd(t.node[i]), ok);
    out.node.push_back(hvm_new_pair(a, b));
  }
  for (hvm_pair r : t.resid) {
    hvm_port a = port(hvm_get_fst(r), ok);
    hvm_port b = port(hvm_get_snd(r), ok);
    out.rbag.push_back(hvm_new_pair(a, b));
  }
  out.vars = nv;
  if (!ok || nv > vars_max) return false;
  for (u32 v = 0; v < t.vars; ++v)
    if (vmap[v] != UINT32_MAX && seen[v] != 2) return false;
  return true;
}

}  // namespace

Book specialize_book(const Book& book, uint32_t node_budget, uint32_t vars_max, std::vector<SpecCounts>& counts) {
  Book out;
  out.defs.resize(book.defs.size());
  counts.assign(book.defs.size(), SpecCounts{});
  for (size_t f = 0; f < book.defs.size(); ++f) {
    const Def& d = book.defs[f];
    out.defs[f] = d;
    Tmpl t;
    t.node = d.node;
    t.alive.assign(d.node.size(), 1);
    t.live = (u32)d.node.size();
    t.root = d.root;
    t.vars = d.vars;
    // the redexes in reverse so the first is looked at first
    t.work.assign(d.rbag.rbegin(), d.rbag.rend());
    SpecCounts s;
    Def nd;
    const u32 budget = std::max<u32>(node_budget, (u32)d.node.size());
    if (!reduce(t, book, budget, s) || (s.call == 0 && s.anni == 0)) continue;
    std::reverse(t.resid.begin(), t.resid.end());
    if (!compact(t, d, vars_max, nd) || nd.node.size() > std::max<size_t>(node_budget, d.node.size())) continue;
    out.defs[f] = std::move(nd);
    s.bytes += (u32)canonical_call_bytes(d);   // the expansion's own bytes, as unspecialised
    counts[f] = s;
  }
  return out;
}

}  // namespace hvm
