"""Synthetic workloads for the evaluator, written as HVM2 core programs.

Each generator returns `.hvm` text plus the closed-form answer the program
must produce.  Inputs come from SplitMix64(seed) with `key = x >> 40` (24 bits).
Key trees are literal perfect binary trees of CON nodes with NUM leaves, split
into defs of at most 5 levels (31 CON nodes) so that every def stays small;
REFs between those defs are expanded lazily by CALL as the program walks them.

Programs (all names and encodings here are this project's own):

* `sum_tree(d)`     — sum(d) = d == 0 ? 1 : sum(d-1) + sum(d-1), the two
                      recursive calls as `&!` redexes; result 2^d mod 2^24.
* `bitonic(k, seed)`— bitonic sort of 2^k keys held as a depth-k tree:
                      sort/flow/down/warp/swap over CON pairs, compare by OPER
                      (`>`, `^`) and branch by SWIT; result: the sorted tree.
* `radix(k, seed)`  — radix sort of 2^k distinct keys through a binary trie of
                      depth 24 (one level per key bit): every key becomes a
                      24-level path, paths are merged pairwise up the key tree,
                      the trie is flattened into an array (Empty/Single/Concat)
                      and reduced to the checksum Σ sorted[i]·i mod 2^24.
* `church(n, m)`    — 2^n independent iterations, each applying the Church
                      numeral 2^m (built by multiplying numerals) to a u24
                      increment starting at 0, summed: result 2^(n+m) mod 2^24.
* `radix_lsd(k, seed)` — LSD radix sort of 2^k uniform u24 keys (duplicates
                      kept), one bit per pass over 24 passes: the keys become a
                      rope (Empty/Single/Concat, Scott-encoded) over the depth-k
                      key tree; each pass is a tree map of a key classifier
                      (a closure copied at every node) with a stable merge of
                      the zeros and ones ropes; then a parallel checksum
                      Σ sorted[i]·i mod 2^24.
* `church_loop(n, m)` — an outer loop of 2^n iterations; iteration i applies
                      the Church numeral 2^m to a u24 increment starting at i
                      and adds the result to the loop's accumulator:
                      Σ_i (i + 2^m) mod 2^24.
"""
from __future__ import annotations

from typing import Iterator, List, Tuple

MASK64 = (1 << 64) - 1
U24 = 1 << 24


def splitmix64(seed: int) -> Iterator[int]:
    x = seed & MASK64
    while True:
        x = (x + 0x9E3779B97F4A7C15) & MASK64
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        yield z ^ (z >> 31)


def keys_uniform(n: int, seed: int) -> List[int]:
    g = splitmix64(seed)
    return [next(g) >> 40 for _ in range(n)]


def keys_distinct(n: int, seed: int) -> List[int]:
    g = splitmix64(seed)
    seen, out = set(), []
    while len(out) < n:
        k = next(g) >> 40
        if k not in seen:
            seen.add(k)
            out.append(k)
    return out


# ---------------------------------------------------------------------------
# Literal key trees
# ---------------------------------------------------------------------------

def key_tree_defs(keys: List[int], name: str, chunk: int = 5) -> Tuple[str, List[str]]:
    """Defs for a perfect tree over `keys` (len a power of two).  Returns the
    top def name and the def lines.  Each def spans at most `chunk` levels."""
    n = len(keys)
    depth = n.bit_length() - 1
    assert 1 << depth == n, "key count must be a power of two"
    lines: List[str] = []
    # level 0: the leaves are numbers
    items = [str(k) for k in keys]
    level = 0
    while True:
        remaining = depth - level
        span = min(chunk, remaining)
        group = 1 << span
        nxt = []
        for gi in range(len(items) // group):
            sub = items[gi * group:(gi + 1) * group]
            while len(sub) > 1:
                sub = [f"({sub[i]} {sub[i + 1]})" for i in range(0, len(sub), 2)]
            if len(items) // group == 1:
                dname = name
            else:
                dname = f"{name}_{level + span}_{gi}"
            lines.append(f"@{dname} = {sub[0]}")
            nxt.append(f"@{dname}")
        items = nxt
        level += span
        if len(items) == 1:
            return name, lines


def flatten_tree(t) -> List[int]:
    if isinstance(t, int):
        return [t]
    return flatten_tree(t[0]) + flatten_tree(t[1])


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------

def sum_tree(depth: int = 20) -> Tuple[str, int]:
    code = f"""
@main = r & @sum ~ ({depth} r)
@sum = (?((1 @sum_s) r) r)
@sum_s = ({{a b}} r)
  &! @sum ~ (a $([+] $(c r)))
  &! @sum ~ (b c)
"""
    return code, (1 << depth) % U24


def stress(depth: int = 8, length: int = 65536) -> Tuple[str, int]:
    """The reference's stress program (examples/stress/main.hvm: 2^depth
    independent loops of `length` iterations each, their results summed)
    with its two sizes as parameters; the shipped example is (8, 65536).
    Its normal form is 0."""
    code = f"""
@fun = (?((@fun__C0 @fun__C1) a) a)
@fun__C0 = a
  & @loop ~ ({length} a)
@fun__C1 = ({{a b}} d)
  &! @fun ~ (a $([+] $(c d)))
  &! @fun ~ (b c)
@loop = (?((0 @loop__C0) a) a)
@loop__C0 = a
  & @loop ~ a
@main = a
  & @fun ~ ({depth} a)
"""
    return code, 0


def tree_gen(depth: int = 30) -> Tuple[str, str]:
    """A full binary tree of CON nodes with 2^depth leaves `1`, built
    recursively: its normal form alone holds 2^depth - 1 nodes (depth 30 does
    not fit a net of 2^29 slots).  Returns the code and, for depth <= 16, the
    printed normal form (else "")."""
    code = f"""
@main = r & @gen ~ ({depth} r)
@gen = (?((1 @gen_s) r) r)
@gen_s = ({{a b}} (x y))
  &! @gen ~ (a x)
  &! @gen ~ (b y)
"""
    nf = ""
    if depth <= 16:
        nf = "1"
        for _ in range(depth):
            nf = f"({nf} {nf})"
    return code, nf


def tree_gen_kinds(depth: int) -> dict:
    """Interaction counts of tree_gen(depth) by kind, as closed forms of the
    depth (ITRS = 13 * 2^depth - 7); tests pin them against the oracle."""
    n = 1 << depth
    k = {"CALL": 3 * n - 1, "CALL_COPY": 0, "VOID": 2 * n - 1, "ERAS": n - 1, "ANNI": 5 * n - 3,
         "COMM": 0, "OPER1": 0, "OPER2": 0, "SWIT0": n, "SWITN": n - 1}
    return k


BITONIC_DEFS = """
@sort = (?((@sort_z @sort_s) (s (t r))) (s (t r)))
@sort_z = (* (t t))
@sort_s = ({da {db dc}} (s ((l u) r)))
  & da ~ $([+1] d)
  & @flow ~ (d (s ((l2 u2) r)))
  &! @sort ~ (db (0 (l l2)))
  &! @sort ~ (dc (1 (u u2)))
@flow = (?((@flow_z @flow_s) (s (t r))) (s (t r)))
@flow_z = (* (t t))
@flow_s = ({da db} ({sa sb} ((a b) r)))
  & db ~ $([+1] d)
  & @warp ~ (da (sa (a (b w))))
  & @down ~ (d (sb (w r)))
@down = (?((@down_z @down_s) (s (t r))) (s (t r)))
@down_z = (* (t t))
@down_s = ({da db} ({sa sb} ((a b) (x y))))
  &! @flow ~ (da (sa (a x)))
  &! @flow ~ (db (sb (b y)))
@warp = (?((@warp_z @warp_s) (s (a (b r)))) (s (a (b r))))
@warp_z = (s ({a1 a2} ({b1 b2} r)))
  & a1 ~ $([>] $(b1 c))
  & c ~ $([^] $(s x))
  & @swap ~ (x (a2 (b2 r)))
@warp_s = ({da db} ({sa sb} ((al ar) ((bl br) ((w0 w2) (w1 w3))))))
  &! @warp ~ (da (sa (al (bl (w0 w1)))))
  &! @warp ~ (db (sb (ar (br (w2 w3)))))
@swap = (?((@swap_z @swap_s) (a (b r))) (a (b r)))
@swap_z = (a (b (a b)))
@swap_s = (* (a (b (b a))))
"""


def bitonic(k: int = 16, seed: int = 0, chunk: int = 5) -> Tuple[str, List[int]]:
    keys = keys_uniform(1 << k, seed)
    top, lines = key_tree_defs(keys, "keys", chunk)
    code = f"@main = r & @sort ~ ({k} (0 (@{top} r)))\n" + BITONIC_DEFS + "\n".join(lines) + "\n"
    return code, sorted(keys)


RADIX_DEFS = """
@Free = (f (* (* f)))
@Used = (* (u (* u)))
@Empty = (e (* (* e)))
@merge = (x (y r)) & x ~ (@m_f (@m_u (@m_n (y r))))
@m_f = (y y)
@m_u = (* @Used)
@m_n = (a (b (y r))) & y ~ (@mn_f (@mn_u (@mn_n (a (b r)))))
@mn_f = (a (b (* (* ((a (b x)) x)))))
@mn_u = (* (* @Used))
@mn_n = (c (d (a (b (* (* ((m1 (m2 x)) x)))))))
  &! @merge ~ (a (c m1))
  &! @merge ~ (b (d m2))
@radix = (?((@rad_z @rad_s) (k r)) (k r))
@rad_z = (* @Used)
@rad_s = ({n1 n2} ({k1 k2} r))
  & k1 ~ $([>>] $(n1 sh))
  & sh ~ $([&] $(1 bit))
  & @radix ~ (n2 (k2 sub))
  & bit ~ ?((@rs_0 @rs_1) (sub r))
@rs_0 = (s (* (* ((s (@Free x)) x))))
@rs_1 = (* (s (* (* ((@Free (s x)) x)))))
@to_map = (?((@tm_z @tm_s) (t r)) (t r))
@tm_z = (k r) & @radix ~ (24 (k r))
@tm_s = ({d1 d2} ((l u) r))
  &! @to_map ~ (d1 (l m1))
  &! @to_map ~ (d2 (u m2))
  & @merge ~ (m1 (m2 r))
@to_arr = (x (p r)) & x ~ (@ta_f (@ta_u (@ta_n (p r))))
@ta_f = (* @Empty)
@ta_u = (p (* ((p x) (* x))))
@ta_n = (l (u ({p1 p2} (* (* ((a (b x)) x))))))
  & p1 ~ $([*2] q0)
  & p2 ~ $([*2] $([+1] q1))
  &! @to_arr ~ (l (q0 a))
  &! @to_arr ~ (u (q1 b))
@cks = (x r) & x ~ (@ck_e (@ck_s (@ck_c r)))
@ck_e = (0 (0 0))
@ck_s = (k (k (0 1)))
@ck_c = (a (b (k (s n))))
  &! @cks ~ (a (ka (sa {na1 na2})))
  &! @cks ~ (b ({kb1 kb2} (sb nb)))
  & ka ~ $([+] $(kb1 k))
  & na1 ~ $([+] $(nb n))
  & kb2 ~ $([*] $(na2 t))
  & sa ~ $([+] $(sb v))
  & v ~ $([+] $(t s))
"""


def radix_checksum(keys: List[int]) -> int:
    return sum(k * i for i, k in enumerate(sorted(set(keys)))) % U24


def radix(k: int = 18, seed: int = 0) -> Tuple[str, int]:
    keys = keys_distinct(1 << k, seed)
    top, lines = key_tree_defs(keys, "keys")
    code = (f"@main = s & @to_map ~ ({k} (@{top} m)) & @to_arr ~ (m (0 a)) & @cks ~ (a (* (s *)))\n"
            + RADIX_DEFS + "\n".join(lines) + "\n")
    return code, radix_checksum(keys)


# LSD radix sort over ropes.  A rope is Scott-encoded: Empty = λe λs λc. e,
# Single k = λe λs λc. s k, Concat l x = λe λs λc. c l x.  `@cat` joins two
# ropes and drops Empty sides (so a rope keeps exactly one Single per key);
# `@split f t` maps the classifier f (key -> 0/1) over t and merges the
# results into the pair (zeros, ones), keeping the order within each side —
# the stable partition of one LSD pass; `@bitf b` is the classifier of bit b.
# `@lsd n b t` runs n passes from bit b; `@rope d t` turns a depth-d key tree
# into a rope; `@cks t` returns (count (sum w)) with w = Σ key_j * j, combined
# over the two halves of a Concat as w1 + w2 + sum2 * count1.
LSD_DEFS = """
@E = (e (* (* e)))
@S = (k (* ((k r) (* r))))
@C = (l (x (* (* ((l (x r)) r)))))
@cat = (x (y r)) & x ~ (@ca_e (@ca_s (@ca_c (y r))))
@ca_e = (y y)
@ca_s = (k (y r)) & y ~ (@cb_e (@cb_s (@cb_c (k r))))
@cb_e = (k r) & @S ~ (k r)
@cb_s = (j (k r)) & @S ~ (k a) & @S ~ (j b) & @C ~ (a (b r))
@cb_c = (l (x (k r))) & @S ~ (k a) & @C ~ (l (x b)) & @C ~ (a (b r))
@ca_c = (l (x (y r))) & y ~ (@cc_e (@cc_s (@cc_c (l (x r)))))
@cc_e = (l (x r)) & @C ~ (l (x r))
@cc_s = (j (l (x r))) & @C ~ (l (x a)) & @S ~ (j b) & @C ~ (a (b r))
@cc_c = (l2 (x2 (l (x r)))) & @C ~ (l (x a)) & @C ~ (l2 (x2 b)) & @C ~ (a (b r))
@split = (f (t r)) & t ~ (@sp_e (@sp_s (@sp_c (f r))))
@sp_e = (* (@E @E))
@sp_s = ({k1 k2} (f r)) & f ~ (k1 bit) & bit ~ ?((@sp_0 @sp_1) (k2 r))
@sp_0 = (k (s @E)) & @S ~ (k s)
@sp_1 = (* (k (@E s))) & @S ~ (k s)
@sp_c = (l (x ({f1 f2} (z o))))
  &! @split ~ (f1 (l (zl ol)))
  &! @split ~ (f2 (x (zx ox)))
  & @cat ~ (zl (zx z))
  & @cat ~ (ol (ox o))
@bitf = (b (k r)) & k ~ $([>>] $(b sh)) & sh ~ $([&] $(1 r))
@lsd = (?((@lsd_z @lsd_s) (b (t r))) (b (t r)))
@lsd_z = (* (t t))
@lsd_s = (n ({b1 b2} (t r)))
  & @bitf ~ (b1 f)
  & @split ~ (f (t (z o)))
  & @cat ~ (z (o t2))
  & b2 ~ $([+1] b3)
  & @lsd ~ (n (b3 (t2 r)))
@rope = (?((@ro_z @ro_s) (t r)) (t r))
@ro_z = (k r) & @S ~ (k r)
@ro_s = ({d1 d2} ((l u) r))
  &! @rope ~ (d1 (l a))
  &! @rope ~ (d2 (u b))
  & @C ~ (a (b r))
@cks = (t r) & t ~ (@ck_e (@ck_s (@ck_c r)))
@ck_e = (0 (0 0))
@ck_s = (k (1 (k 0)))
@ck_c = (l (x (c (s w))))
  &! @cks ~ (l ({c1a c1b} (s1 w1)))
  &! @cks ~ (x (c2 ({s2a s2b} w2)))
  & c1a ~ $([+] $(c2 c))
  & s1 ~ $([+] $(s2a s))
  & s2b ~ $([*] $(c1b t))
  & w1 ~ $([+] $(w2 u))
  & u ~ $([+] $(t w))
"""


def lsd_checksum(keys: List[int]) -> int:
    return sum(k * i for i, k in enumerate(sorted(keys))) % U24


def radix_lsd(k: int = 18, seed: int = 0, bits: int = 24) -> Tuple[str, int]:
    """LSD radix sort of 2^k uniform u24 keys (with duplicates), `bits`
    one-bit passes; result Σ sorted[i]·i mod 2^24."""
    keys = keys_uniform(1 << k, seed)
    top, lines = key_tree_defs(keys, "keys")
    code = (f"@main = w & @rope ~ ({k} (@{top} t)) & @lsd ~ ({bits} (0 (t t2))) & @cks ~ (t2 (* (* w)))\n"
            + LSD_DEFS + "\n".join(lines) + "\n")
    if bits == 24:
        return code, lsd_checksum(keys)
    part = sorted(keys, key=lambda x: x & ((1 << bits) - 1))   # stable: by the low bits
    return code, sum(x * i for i, x in enumerate(part)) % U24


def church_numeral_defs(m: int) -> Tuple[List[str], str]:
    """Defs of the Church numeral 2^m, built by multiplying numerals (2^1,
    2^2, 2^4, ... by squaring, then the binary decomposition of m); returns
    the lines and the numeral's def name."""
    assert m >= 1
    lines = ["@inc = ($([+1] r) r)",
             "@c1 = ({(x y) (y r)} (x r))",
             "@mul = (f (g (h r))) & g ~ (h i) & f ~ (i r)"]
    have = {1: "c1"}
    e = 1
    while e * 2 <= m:
        lines.append(f"@c{e * 2} = r & @mul ~ (@c{e} (@c{e} r))")
        have[e * 2] = f"c{e * 2}"
        e *= 2
    parts = [p for p in sorted(have) if m & p]
    if sum(parts) != m:
        raise ValueError("m must be a sum of distinct powers of two")
    cur = have[parts[0]]
    for p in parts[1:]:
        nxt = f"{cur}x{p}"
        lines.append(f"@{nxt} = r & @mul ~ (@{cur} (@{have[p]} r))")
        cur = nxt
    return lines, cur


def church_loop(n: int = 12, m: int = 12) -> Tuple[str, int]:
    """Outer loop of 2^n iterations: iteration i applies the Church numeral
    2^m to the increment starting at i; the loop sums the results."""
    lines, cur = church_numeral_defs(m)
    lines += ["@loop = (?((@loop_z @loop_s) (a r)) (a r))",
              "@loop_z = (a a)",
              f"@loop_s = ({{i1 i2}} (a r)) & @{cur} ~ (@inc (i1 v)) & a ~ $([+] $(v a2)) & @loop ~ (i2 (a2 r))",
              f"@main = r & @loop ~ ({1 << n} (0 r))"]
    N = 1 << n
    return "\n".join(lines) + "\n", (N * (N - 1) // 2 + N * (1 << m)) % U24


def church(n: int = 12, m: int = 12) -> Tuple[str, int]:
    """2^n iterations of (Church 2^m) applied to x+1 from 0, summed."""
    assert m >= 1
    # numerals 2^1, 2^2, 2^4, ... by squaring, then 2^m as a product of them
    lines = ["@inc = ($([+1] r) r)",
             "@c1 = ({(x y) (y r)} (x r))",
             "@mul = (f (g (h r))) & g ~ (h i) & f ~ (i r)"]
    have = {1: "c1"}
    e = 1
    while e * 2 <= m:
        lines.append(f"@c{e * 2} = r & @mul ~ (@c{e} (@c{e} r))")
        have[e * 2] = f"c{e * 2}"
        e *= 2
    # binary decomposition of m
    parts = [p for p in sorted(have) if m & p]
    if sum(parts) != m:
        raise ValueError("m must be a sum of distinct powers of two")
    cur = have[parts[0]]
    for p in parts[1:]:
        nxt = f"{cur}x{p}"
        lines.append(f"@{nxt} = r & @mul ~ (@{cur} (@{have[p]} r))")
        cur = nxt
    lines += [f"@leaf = r & @{cur} ~ (@inc (0 r))",
              "@fan = (?((@leaf @fan_s) r) r)",
              "@fan_s = ({a b} r) &! @fan ~ (a $([+] $(c r))) &! @fan ~ (b c)",
              f"@main = r & @fan ~ ({n} r)"]
    return "\n".join(lines) + "\n", (1 << (n + m)) % U24


def parse_tree_text(s: str):
    """Parses a printed tree of CON pairs and u24 numbers into nested tuples."""
    pos = 0

    def walk():
        nonlocal pos
        if s[pos] == "(":
            pos += 1
            a = walk()
            pos += 1  # space
            b = walk()
            pos += 1  # ')'
            return (a, b)
        j = pos
        while j < len(s) and s[j].isdigit():
            j += 1
        v = int(s[pos:j])
        pos = j
        return v

    return walk()


SUM_DEFS = """
@sum = (?((1 @sum_s) r) r)
@sum_s = ({a b} r)
  &! @sum ~ (a $([+] $(c r)))
  &! @sum ~ (b c)
"""


def batch(seed: int = 0, n: int = 256, chunk: int = 8):
    """Batched workload: n independent nets drawn from `seed`; the first half
    are tree sums of depth U[10, 22], the second half bitonic sorts of 2^k keys
    with k = U[8, 14].  Returns (code, nets) where code is one book holding
    every net (entry def `@net<i>`) and nets[i] = (generator, params, expected,
    standalone_code): the same net as a program of its own, whose run-c result
    is the per-net parity target."""
    g = splitmix64(seed ^ 0x5EED_BA7C)
    nets = []
    body = ["@main = *", SUM_DEFS.strip(), BITONIC_DEFS.strip()]
    for i in range(n):
        if i < n // 2:
            d = 10 + next(g) % 13
            code_i, expect = sum_tree(d)
            nets.append(("sum_tree", {"depth": d}, expect, code_i))
            body.append(f"@net{i} = r & @sum ~ ({d} r)")
        else:
            k = 8 + next(g) % 7
            ks = next(g) & 0xFFFFFFFF
            code_i, expect = bitonic(k, ks, chunk)
            nets.append(("bitonic", {"k": k, "seed": ks, "chunk": chunk}, expect, code_i))
            top, lines = key_tree_defs(keys_uniform(1 << k, ks), f"k{i}", chunk)
            body.append(f"@net{i} = r & @sort ~ ({k} (0 (@{top} r)))")
            body.extend(lines)
    return "\n".join(body) + "\n", nets


def def_fids(code: str) -> dict:
    """fid of every def, as the builder assigns them (main first, then byte order)."""
    names = sorted({line.split("=", 1)[0].strip()[1:] for line in code.splitlines()
                    if line.startswith("@") and "=" in line})
    order = ["main"] + [x for x in names if x != "main"]
    return {name: i for i, name in enumerate(order)}


def batch_estimate(gen: str, params: dict) -> int:
    """Estimated peak live nodes of one net of the batched workload, used for
    admission: a tree sum keeps O(2^d) partial sums alive when evaluated
    breadth-first; a bitonic sort holds its key tree plus one merge level."""
    if gen == "sum_tree":
        return 8 << params["depth"]
    k = params["k"]
    return 16 * k * (1 << k)


CONFIGS = {
    "sum_tree_d20": lambda: sum_tree(20),
    "bitonic_2e16": lambda: bitonic(16, 0),
    "radix_2e18": lambda: radix(18, 0),
    "church_2e12x2e12": lambda: church(12, 12),
    "radix_lsd_2e18": lambda: radix_lsd(18, 0),
    "church_loop_2e12x2e12": lambda: church_loop(12, 12),
}
