/* hvm_oracle.c — CPU oracle for the HVM2 GPU path.  TEST INFRASTRUCTURE ONLY:
 * used as the checker by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg; never part of the product.  See hvm_oracle.h for the
 * reference lines each part restates and for how the oracle is pinned. */
#define _GNU_SOURCE
#include "hvm_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../lambdavm-rust_amd/csrc/hvm_core.h"

typedef uint32_t Port;
typedef uint64_t Pair;

/* Address space of the reference: 2^29 nodes and vars (src/hvm.c:128-129);
 * ROOT is var 0x1FFFFFFF, so the var buffer spans the full range.  Both
 * buffers are calloc'd and committed lazily by the OS. */
#define G_LEN (1ull << 29)

typedef struct {
  Pair* node;
  Port* vars;
  /* high- and low-priority redex stacks (src/hvm.c:589-621) */
  Pair* hbag; size_t hlen, hcap;
  Pair* rbag; size_t rlen, rcap;
  /* allocator cursors and scratch (src/hvm.c:177-186) */
  uint64_t nput, vput;
  uint32_t* nloc; uint32_t* vloc; size_t lcap;
  /* book */
  uint32_t ndefs;
  struct ODef { char name[256]; int safe; uint32_t rlen, nlen, vlen; Port root; Pair* rbag; Pair* node; } *defs;
  hvm_oracle_stats* st;
  uint64_t live_nodes;
  int oom;
  /* optional: expansions per fid (hvm_oracle_run_calls) */
  uint64_t* calls; uint32_t ncalls;
} ONet;

/* ----------------------------------------------------------------------- */
/* bags (src/hvm.c:589-621)                                                 */
/* ----------------------------------------------------------------------- */

static int grow(Pair** buf, size_t* cap) {
  size_t nc = *cap ? *cap * 2 : 1024;
  Pair* nb = (Pair*)realloc(*buf, nc * sizeof(Pair));
  if (!nb) return 0;
  *buf = nb; *cap = nc;
  return 1;
}

static void push_redex(ONet* n, Pair r) {
  if (hvm_is_high_priority(hvm_get_rule(hvm_get_fst(r), hvm_get_snd(r)))) {
    if (n->hlen == n->hcap && !grow(&n->hbag, &n->hcap)) { n->oom = 1; return; }
    n->hbag[n->hlen++] = r;
  } else {
    if (n->rlen == n->rcap && !grow(&n->rbag, &n->rcap)) { n->oom = 1; return; }
    n->rbag[n->rlen++] = r;
  }
}

static Pair pop_redex(ONet* n) {
  if (n->hlen) return n->hbag[--n->hlen];
  if (n->rlen) return n->rbag[--n->rlen];
  return 0;
}

/* ----------------------------------------------------------------------- */
/* net accessors (src/hvm.c:652-694)                                        */
/* ----------------------------------------------------------------------- */

static inline void node_create(ONet* n, uint32_t loc, Pair v) { n->node[loc] = v; n->live_nodes++;
  if (n->live_nodes > n->st->peak_nodes) n->st->peak_nodes = n->live_nodes; }
static inline Pair node_take(ONet* n, uint32_t loc) { Pair v = n->node[loc]; n->node[loc] = 0; if (v) n->live_nodes--; return v; }
static inline void vars_create(ONet* n, uint32_t var, Port v) { n->vars[var] = v; }
static inline Port vars_load(ONet* n, uint32_t var) { return n->vars[var]; }
static inline Port vars_exchange(ONet* n, uint32_t var, Port v) { Port o = n->vars[var]; n->vars[var] = v; return o; }
static inline Port vars_take(ONet* n, uint32_t var) { return vars_exchange(n, var, 0); }

/* ----------------------------------------------------------------------- */
/* allocators (src/hvm.c:739-780), single slice covering the whole space    */
/* ----------------------------------------------------------------------- */

static int ensure_loc(ONet* n, size_t need) {
  if (need <= n->lcap) return 1;
  size_t nc = need < 4096 ? 4096 : need;
  uint32_t* a = (uint32_t*)realloc(n->nloc, nc * 4);
  if (!a) return 0;
  n->nloc = a;
  uint32_t* b = (uint32_t*)realloc(n->vloc, nc * 4);
  if (!b) return 0;
  n->vloc = b;
  n->lcap = nc;
  return 1;
}

static int node_alloc(ONet* n, uint32_t num) {
  uint32_t got = 0;
  uint64_t lps = 0;
  while (got < num) {
    uint32_t lc = (uint32_t)(n->nput % G_LEN);
    Pair elem = n->node[lc];
    n->nput += 1;
    if (lc > 0 && elem == 0) n->nloc[got++] = lc;
    if (++lps >= G_LEN) { n->oom = 1; return 0; }
  }
  return 1;
}

static int vars_alloc(ONet* n, uint32_t num) {
  uint32_t got = 0;
  uint64_t lps = 0;
  while (got < num) {
    uint32_t lc = (uint32_t)(n->vput % G_LEN);
    Port elem = n->vars[lc];
    n->vput += 1;
    if (lc > 0 && elem == 0) n->vloc[got++] = lc;
    if (++lps >= G_LEN) { n->oom = 1; return 0; }
  }
  return 1;
}

static int get_resources(ONet* n, uint32_t need_node, uint32_t need_vars) {
  if (!ensure_loc(n, need_node > need_vars ? need_node : need_vars)) { n->oom = 1; return 0; }
  return node_alloc(n, need_node) && vars_alloc(n, need_vars);
}

/* ----------------------------------------------------------------------- */
/* linking (src/hvm.c:784-846)                                              */
/* ----------------------------------------------------------------------- */

static Port peek(ONet* n, Port var) {
  while (hvm_get_tag(var) == HVM_VAR) {
    Port val = vars_load(n, hvm_get_val(var));
    if (val == HVM_NONE || val == 0) break;
    var = val;
  }
  return var;
}

static Port enter(ONet* n, Port var) {
  while (hvm_get_tag(var) == HVM_VAR) {
    Port val = vars_exchange(n, hvm_get_val(var), HVM_NONE);
    if (val == HVM_NONE || val == 0) break;
    vars_take(n, hvm_get_val(var));
    var = val;
  }
  return var;
}

static void link_ports(ONet* n, Port A, Port B) {
  for (;;) {
    if (hvm_get_tag(A) != HVM_VAR && hvm_get_tag(B) == HVM_VAR) { Port x = A; A = B; B = x; }
    if (hvm_get_tag(A) != HVM_VAR) { push_redex(n, hvm_new_pair(A, B)); return; }
    B = enter(n, B);
    Port A_ = vars_exchange(n, hvm_get_val(A), B);
    if (A_ == HVM_NONE) return;
    vars_take(n, hvm_get_val(A));
    A = A_;
  }
}

static inline Port adjust_port(ONet* n, Port p) {
  hvm_tag t = hvm_get_tag(p);
  if (hvm_is_nod(p)) return hvm_new_port(t, n->nloc[hvm_get_val(p)]);
  if (t == HVM_VAR) return hvm_new_port(t, n->vloc[hvm_get_val(p)]);
  return p;
}

static inline Pair adjust_pair(ONet* n, Pair p) {
  return hvm_new_pair(adjust_port(n, hvm_get_fst(p)), adjust_port(n, hvm_get_snd(p)));
}

/* ----------------------------------------------------------------------- */
/* interactions (src/hvm.c:852-1069)                                        */
/* ----------------------------------------------------------------------- */

static void count(ONet* n, int kind, uint64_t bytes) {
  n->st->kind[kind] += 1;
  n->st->bytes += bytes;
}

static int interact_eras(ONet* n, Port a, Port b) {
  if (n->node[hvm_get_val(b)] == 0) return 0;
  Pair B = node_take(n, hvm_get_val(b));
  link_ports(n, a, hvm_get_fst(B));
  link_ports(n, a, hvm_get_snd(B));
  return 1;
}

static int interact_call(ONet* n, Port a, Port b) {
  uint32_t fid = hvm_get_val(a) & 0xFFFFFFFu;
  struct ODef* d = &n->defs[fid];
  if (d->safe && hvm_get_tag(b) == HVM_DUP) {
    if (!interact_eras(n, a, b)) return 0;
    count(n, HVM_ORC_CALL_COPY, 32);
    return 1;
  }
  if (!get_resources(n, d->nlen, d->vlen)) return 0;
  for (uint32_t i = 0; i < d->vlen; ++i) vars_create(n, n->vloc[i], HVM_NONE);
  for (uint32_t i = 0; i < d->nlen; ++i) node_create(n, n->nloc[i], adjust_pair(n, d->node[i]));
  for (uint32_t i = 0; i < d->rlen; ++i) {
    Pair r = adjust_pair(n, d->rbag[i]);
    link_ports(n, hvm_get_fst(r), hvm_get_snd(r));
  }
  link_ports(n, adjust_port(n, d->root), b);
  count(n, HVM_ORC_CALL, 8ull * d->nlen + 4ull * d->vlen + 8ull * (d->rlen + 1));
  if (n->calls && fid < n->ncalls) n->calls[fid] += 1;
  return 1;
}

static int interact_anni(ONet* n, Port a, Port b) {
  if (n->node[hvm_get_val(a)] == 0 || n->node[hvm_get_val(b)] == 0) return 0;
  Pair A = node_take(n, hvm_get_val(a));
  Pair B = node_take(n, hvm_get_val(b));
  link_ports(n, hvm_get_fst(A), hvm_get_fst(B));
  link_ports(n, hvm_get_snd(A), hvm_get_snd(B));
  count(n, HVM_ORC_ANNI, 48);
  return 1;
}

static int interact_comm(ONet* n, Port a, Port b) {
  if (!get_resources(n, 4, 4)) return 0;
  if (n->node[hvm_get_val(a)] == 0 || n->node[hvm_get_val(b)] == 0) return 0;
  Pair A = node_take(n, hvm_get_val(a));
  Pair B = node_take(n, hvm_get_val(b));
  uint32_t* v = n->vloc;
  uint32_t* l = n->nloc;
  for (int i = 0; i < 4; ++i) vars_create(n, v[i], HVM_NONE);
  node_create(n, l[0], hvm_new_pair(hvm_new_port(HVM_VAR, v[0]), hvm_new_port(HVM_VAR, v[1])));
  node_create(n, l[1], hvm_new_pair(hvm_new_port(HVM_VAR, v[2]), hvm_new_port(HVM_VAR, v[3])));
  node_create(n, l[2], hvm_new_pair(hvm_new_port(HVM_VAR, v[0]), hvm_new_port(HVM_VAR, v[2])));
  node_create(n, l[3], hvm_new_pair(hvm_new_port(HVM_VAR, v[1]), hvm_new_port(HVM_VAR, v[3])));
  link_ports(n, hvm_new_port(hvm_get_tag(b), l[0]), hvm_get_fst(A));
  link_ports(n, hvm_new_port(hvm_get_tag(b), l[1]), hvm_get_snd(A));
  link_ports(n, hvm_new_port(hvm_get_tag(a), l[2]), hvm_get_fst(B));
  link_ports(n, hvm_new_port(hvm_get_tag(a), l[3]), hvm_get_snd(B));
  count(n, HVM_ORC_COMM, 112);
  return 1;
}

static int interact_oper(ONet* n, Port a, Port b) {
  if (!get_resources(n, 1, 0)) return 0;
  if (n->node[hvm_get_val(b)] == 0) return 0;
  uint32_t av = hvm_get_val(a);
  Pair B = node_take(n, hvm_get_val(b));
  Port B1 = hvm_get_fst(B);
  Port B2 = enter(n, hvm_get_snd(B));
  if (hvm_get_tag(B1) == HVM_NUM) {
    uint32_t cv = hvm_operate(av, hvm_get_val(B1));
    link_ports(n, hvm_new_port(HVM_NUM, cv), B2);
    count(n, HVM_ORC_OPER1, 32);
  } else {
    node_create(n, n->nloc[0], hvm_new_pair(a, B2));
    link_ports(n, B1, hvm_new_port(HVM_OPR, n->nloc[0]));
    count(n, HVM_ORC_OPER2, 40);
  }
  return 1;
}

static int interact_swit(ONet* n, Port a, Port b) {
  if (!get_resources(n, 2, 0)) return 0;
  if (n->node[hvm_get_val(b)] == 0) return 0;
  uint32_t av = hvm_get_u24(hvm_get_val(a));
  Pair B = node_take(n, hvm_get_val(b));
  Port B1 = hvm_get_fst(B), B2 = hvm_get_snd(B);
  if (av == 0) {
    node_create(n, n->nloc[0], hvm_new_pair(B2, hvm_new_port(HVM_ERA, 0)));
    link_ports(n, hvm_new_port(HVM_CON, n->nloc[0]), B1);
    count(n, HVM_ORC_SWIT0, 32);
  } else {
    node_create(n, n->nloc[0], hvm_new_pair(hvm_new_port(HVM_ERA, 0), hvm_new_port(HVM_CON, n->nloc[1])));
    node_create(n, n->nloc[1], hvm_new_pair(hvm_new_port(HVM_NUM, hvm_new_u24(av - 1)), B2));
    link_ports(n, hvm_new_port(HVM_CON, n->nloc[0]), B1);
    count(n, HVM_ORC_SWITN, 40);
  }
  return 1;
}

/* src/hvm.c:1073-1121 */
static void interact(ONet* n, Pair redex) {
  Port a = hvm_get_fst(redex), b = hvm_get_snd(redex);
  hvm_rule rule = hvm_get_rule(a, b);
  if (hvm_get_tag(a) == HVM_REF && b == HVM_ROOT) {
    rule = HVM_CALL;
  } else if (hvm_should_swap(a, b)) {
    Port x = a; a = b; b = x;
  }
  int ok = 1;
  switch (rule) {
    case HVM_LINK: link_ports(n, a, b); break;
    case HVM_CALL: ok = interact_call(n, a, b); break;
    case HVM_VOID: count(n, HVM_ORC_VOID, 0); break;
    case HVM_ERAS: ok = interact_eras(n, a, b); if (ok) count(n, HVM_ORC_ERAS, 32); break;
    case HVM_ANNI: ok = interact_anni(n, a, b); break;
    case HVM_COMM: ok = interact_comm(n, a, b); break;
    case HVM_OPER: ok = interact_oper(n, a, b); break;
    case HVM_SWIT: ok = interact_swit(n, a, b); break;
  }
  if (!ok) { push_redex(n, redex); n->oom |= 2; return; }
  if (rule != HVM_LINK) n->st->itrs += 1;
}

static void boot_redex(ONet* n, Pair redex) {
  n->vars[hvm_get_val(HVM_ROOT)] = HVM_NONE;
  push_redex(n, redex);
}

static void normalize(ONet* n) {
  n->st->normalizes += 1;
  for (;;) {
    if (n->oom) return;
    Pair r = pop_redex(n);
    if (!r) break;
    interact(n, r);
  }
}

/* src/hvm.c:1233-1244 */
static Port expand(ONet* n, Port port) {
  Port old = vars_load(n, hvm_get_val(HVM_ROOT));
  Port got = peek(n, port);
  while (hvm_get_tag(got) == HVM_REF && !n->oom) {
    boot_redex(n, hvm_new_pair(got, HVM_ROOT));
    normalize(n);
    got = peek(n, vars_load(n, hvm_get_val(HVM_ROOT)));
  }
  vars_create(n, hvm_get_val(HVM_ROOT), old);
  return got;
}

/* src/run.c:67-100 */
typedef struct { int32_t tag; uint32_t args_len; Port args[16]; } Ctr;

static Ctr readback_ctr(ONet* n, Port port) {
  Ctr c;
  c.tag = -1;
  c.args_len = 0;
  Port lam = expand(n, port);
  if (hvm_get_tag(lam) != HVM_CON) return c;
  Pair lam_node = n->node[hvm_get_val(lam)];
  Port app = expand(n, hvm_get_fst(lam_node));
  if (hvm_get_tag(app) != HVM_CON) return c;
  Pair app_node = n->node[hvm_get_val(app)];
  Port arg = expand(n, hvm_get_fst(app_node));
  if (hvm_get_tag(arg) != HVM_NUM) return c;
  c.tag = (int32_t)hvm_get_u24(hvm_get_val(arg));
  for (;;) {
    app = expand(n, hvm_get_snd(app_node));
    if (hvm_get_tag(app) != HVM_CON) break;
    app_node = n->node[hvm_get_val(app)];
    arg = expand(n, hvm_get_fst(app_node));
    if (c.args_len < 16) c.args[c.args_len] = arg;
    c.args_len++;
  }
  return c;
}

/* src/run.c:740-810, without the IO calls (returns 2 if one is requested). */
static int run_io(ONet* n) {
  Port port = HVM_ROOT;
  normalize(n);
  Ctr c = readback_ctr(n, peek(n, port));
  if (c.args_len < 1 || hvm_get_tag(c.args[0]) != HVM_CON) return 0;
  Pair magic = n->node[hvm_get_val(c.args[0])];
  if (hvm_get_val(hvm_get_fst(magic)) != hvm_new_u24(0xD0CA11) ||
      hvm_get_val(hvm_get_snd(magic)) != hvm_new_u24(0xFF1FF1)) return 0;
  if (c.tag == 1 && c.args_len == 4) return 2;
  return 0;
}

/* ----------------------------------------------------------------------- */
/* printer (src/hvm.c:1535-1699)                                            */
/* ----------------------------------------------------------------------- */

typedef struct { char* s; size_t len, cap; } Str;

static void sput(Str* b, const char* fmt, ...) {
  char tmp[96];
  va_list ap;
  va_start(ap, fmt);
  int k = vsnprintf(tmp, sizeof tmp, fmt, ap);
  va_end(ap);
  if (b->len + (size_t)k + 1 > b->cap) {
    size_t nc = b->cap ? b->cap * 2 : 256;
    while (nc < b->len + (size_t)k + 1) nc *= 2;
    b->s = (char*)realloc(b->s, nc);
    b->cap = nc;
  }
  memcpy(b->s + b->len, tmp, (size_t)k + 1);
  b->len += (size_t)k;
}

static const char* op_sym(uint32_t op) {
  static const char* t[] = {"", "", "", "", "+", "-", ":-", "*", "/", ":/", "%", ":%",
                            "=", "!", "<", ">", "&", "|", "^", "<<", ":<<", ">>", ":>>"};
  return op <= 0x16 && op >= 4 ? t[op] : "?";
}

static void print_numb(Str* b, uint32_t w) {
  switch (hvm_get_typ(w)) {
    case HVM_TY_SYM: {
      uint32_t s = hvm_get_sym(w);
      if (s == HVM_TY_U24) sput(b, "[u24]");
      else if (s == HVM_TY_I24) sput(b, "[i24]");
      else if (s == HVM_TY_F24) sput(b, "[f24]");
      else sput(b, "[%s]", op_sym(s));
      break;
    }
    case HVM_TY_U24: sput(b, "%u", hvm_get_u24(w)); break;
    case HVM_TY_I24: sput(b, "%+d", hvm_get_i24(w)); break;
    case HVM_TY_F24: {
      float f = hvm_get_f24(w);
      if (isinf(f)) sput(b, signbit(f) ? "-inf" : "+inf");
      else if (isnan(f)) sput(b, "+NaN");
      else sput(b, "%.7e", (double)f);
      break;
    }
    default: sput(b, "[%s0x%07X]", op_sym(hvm_get_typ(w)), hvm_get_u24(w)); break;
  }
}

static char* print_port(ONet* n, Port port) {
  Str b = {0, 0, 0};
  sput(&b, "%s", "");
  size_t cap = 4096, len = 1;
  Port* stack = (Port*)malloc(cap * sizeof(Port));
  stack[0] = port;
  while (len > 0) {
    Port cur = stack[--len];
    if (len + 4 > cap) { cap *= 2; stack = (Port*)realloc(stack, cap * sizeof(Port)); }
    switch (hvm_get_tag(cur)) {
      case HVM_CON: case HVM_DUP: case HVM_OPR: case HVM_SWI: {
        Pair node = n->node[hvm_get_val(cur)];
        const char* open = hvm_get_tag(cur) == HVM_CON ? "(" : hvm_get_tag(cur) == HVM_DUP ? "{"
                         : hvm_get_tag(cur) == HVM_OPR ? "$(" : "?(";
        char close = hvm_get_tag(cur) == HVM_DUP ? '}' : ')';
        sput(&b, "%s", open);
        stack[len++] = hvm_new_port(HVM_ERA, (uint32_t)close);
        stack[len++] = hvm_get_snd(node);
        stack[len++] = hvm_new_port(HVM_ERA, (uint32_t)' ');
        stack[len++] = hvm_get_fst(node);
        break;
      }
      case HVM_ERA:
        if (hvm_get_val(cur) != 0) sput(&b, "%c", (char)hvm_get_val(cur));
        else sput(&b, "*");
        break;
      case HVM_VAR: {
        Port got = vars_load(n, hvm_get_val(cur));
        if (got != HVM_NONE) stack[len++] = got;
        else sput(&b, "x%x", hvm_get_val(cur));
        break;
      }
      case HVM_NUM: print_numb(&b, hvm_get_val(cur)); break;
      case HVM_REF: sput(&b, "@%s", n->defs[hvm_get_val(cur) & 0xFFFFFFFu].name); break;
    }
  }
  free(stack);
  return b.s;
}

/* ----------------------------------------------------------------------- */
/* entry (src/hvm.c:1743-1791)                                              */
/* ----------------------------------------------------------------------- */

static int load_book(ONet* n, const uint32_t* buf, size_t n_words) {
  size_t i = 0;
#define NEED(k) do { if (n_words != (size_t)-1 && i + (k) > n_words) return 0; } while (0)
  NEED(1);
  n->ndefs = buf[i++];
  if (n->ndefs > HVM_MAX_DEFS) return 0;
  n->defs = (struct ODef*)calloc(n->ndefs ? n->ndefs : 1, sizeof(struct ODef));
  for (uint32_t k = 0; k < n->ndefs; ++k) {
    NEED(70);
    uint32_t fid = buf[i++];
    if (fid >= n->ndefs) return 0;
    struct ODef* d = &n->defs[fid];
    memcpy(d->name, buf + i, 256);
    d->name[255] = 0;
    i += 64;
    d->safe = buf[i++] != 0;
    d->rlen = buf[i++];
    d->nlen = buf[i++];
    d->vlen = buf[i++];
    if (d->rlen > HVM_DEF_RBAG_LEN || d->nlen > HVM_DEF_NODE_LEN) return 0;
    d->root = buf[i++];
    NEED(2 * ((size_t)d->rlen + d->nlen));
    d->rbag = (Pair*)malloc((d->rlen + 1) * sizeof(Pair));
    d->node = (Pair*)malloc((d->nlen + 1) * sizeof(Pair));
    memcpy(d->rbag, buf + i, 8ull * d->rlen); i += 2ull * d->rlen;
    memcpy(d->node, buf + i, 8ull * d->nlen); i += 2ull * d->nlen;
  }
#undef NEED
  return 1;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void free_net(ONet* n) {
  if (n->defs) {
    for (uint32_t k = 0; k < n->ndefs; ++k) { free(n->defs[k].rbag); free(n->defs[k].node); }
    free(n->defs);
  }
  free(n->node); free(n->vars); free(n->hbag); free(n->rbag); free(n->nloc); free(n->vloc);
}

int hvm_oracle_run_calls(const uint32_t* book_buf, size_t n_words, char** result, hvm_oracle_stats* st,
                         uint64_t* calls, uint32_t ncalls) {
  hvm_oracle_stats local;
  if (!st) st = &local;
  memset(st, 0, sizeof *st);
  if (result) *result = NULL;
  ONet n;
  memset(&n, 0, sizeof n);
  n.st = st;
  n.calls = calls;
  n.ncalls = calls ? ncalls : 0;
  if (calls) memset(calls, 0, sizeof(uint64_t) * ncalls);
  n.nput = 1;
  n.vput = 1;
  if (!book_buf || !load_book(&n, book_buf, n_words) || n.ndefs == 0) { free_net(&n); return 1; }
  n.node = (Pair*)calloc(G_LEN, sizeof(Pair));
  n.vars = (Port*)calloc(G_LEN, sizeof(Port));
  if (!n.node || !n.vars || !ensure_loc(&n, 4096)) { free_net(&n); return 3; }
  double t0 = now_s();
  boot_redex(&n, hvm_new_pair(hvm_new_port(HVM_REF, 0), HVM_ROOT));
  int rc = run_io(&n);
  if (n.oom) { free_net(&n); return 3; }
  if (rc == 2) { free_net(&n); return 2; }
  char* text = print_port(&n, enter(&n, HVM_ROOT));
  st->seconds = now_s() - t0;
  if (result) *result = text; else free(text);
  free_net(&n);
  return 0;
}

int hvm_oracle_run(const uint32_t* book_buf, size_t n_words, char** result, hvm_oracle_stats* st) {
  return hvm_oracle_run_calls(book_buf, n_words, result, st, NULL, 0);
}

void hvm_oracle_free(void* p) { free(p); }
