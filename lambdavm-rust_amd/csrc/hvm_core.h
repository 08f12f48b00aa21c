// hvm_core.h — HVM2 value encodings, interaction-rule table and numeric ALU.
//
// Shared by the host front end (compiled with g++) and the HIP evaluator
// (compiled with hipcc for gfx950).  Every encoding here is bit-compatible with
// the reference so that book buffers and device nets are interchangeable:
//
//   Port = val << 3 | tag           (3-bit tag, 29-bit val)   src/hvm.rs:126-137, src/hvm.c:197-207
//   Pair = snd << 32 | fst          (two ports in a u64)       src/hvm.rs:176-186, src/hvm.c:211-221
//   Numb = payload << 5 | typ       (5-bit typ, 24-bit value)  src/hvm.rs:224-281, src/hvm.c:370-425
//
// The rule table and priorities follow src/hvm.c:319-348 (src/hvm.rs:147-168);
// the ALU (`hvm_operate`) follows src/hvm.c:427-582 (src/hvm.rs:286-417).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HVM_HD __host__ __device__ __forceinline__
// rarely taken code kept out of line: its constants stay out of the caller's
// registers (the f24 operators' libm constants would otherwise be spilled)
#define HVM_HD_COLD static __host__ __device__ __attribute__((noinline))
#else
#include <math.h>
#include <string.h>
#define HVM_HD static inline
#define HVM_HD_COLD static
#endif

typedef uint8_t  hvm_tag;
typedef uint32_t hvm_val;
typedef uint32_t hvm_port;
typedef uint64_t hvm_pair;
typedef uint32_t hvm_numb;
typedef uint8_t  hvm_rule;

// Port tags (src/hvm.c:65-72)
enum {
  HVM_VAR = 0, HVM_REF = 1, HVM_ERA = 2, HVM_NUM = 3,
  HVM_CON = 4, HVM_DUP = 5, HVM_OPR = 6, HVM_SWI = 7,
};

// Interaction rules (src/hvm.c:75-82)
enum {
  HVM_LINK = 0, HVM_CALL = 1, HVM_VOID = 2, HVM_ERAS = 3,
  HVM_ANNI = 4, HVM_COMM = 5, HVM_OPER = 6, HVM_SWIT = 7,
};

// Numeric types and operators (src/hvm.c:89-111)
enum {
  HVM_TY_SYM = 0x00, HVM_TY_U24 = 0x01, HVM_TY_I24 = 0x02, HVM_TY_F24 = 0x03,
  HVM_OP_ADD = 0x04, HVM_OP_SUB = 0x05, HVM_FP_SUB = 0x06, HVM_OP_MUL = 0x07,
  HVM_OP_DIV = 0x08, HVM_FP_DIV = 0x09, HVM_OP_REM = 0x0A, HVM_FP_REM = 0x0B,
  HVM_OP_EQ  = 0x0C, HVM_OP_NEQ = 0x0D, HVM_OP_LT  = 0x0E, HVM_OP_GT  = 0x0F,
  HVM_OP_AND = 0x10, HVM_OP_OR  = 0x11, HVM_OP_XOR = 0x12, HVM_OP_SHL = 0x13,
  HVM_FP_SHL = 0x14, HVM_OP_SHR = 0x15, HVM_FP_SHR = 0x16,
};

// Special ports (src/hvm.c:114-116)
#define HVM_FREE 0x00000000u
#define HVM_ROOT 0xFFFFFFF8u
#define HVM_NONE 0xFFFFFFFFu

// Book limits enforced by the loader (src/hvm.c:135-136, 159-164).
#define HVM_MAX_DEFS     0x4000u
#define HVM_DEF_NODE_LEN 0xFFFu
#define HVM_DEF_RBAG_LEN 0xFFFu

// ---------------------------------------------------------------------------
// Port / Pair
// ---------------------------------------------------------------------------

HVM_HD hvm_port hvm_new_port(hvm_tag tag, hvm_val val) { return (val << 3) | tag; }
HVM_HD hvm_tag  hvm_get_tag(hvm_port p) { return (hvm_tag)(p & 7u); }
HVM_HD hvm_val  hvm_get_val(hvm_port p) { return p >> 3; }
HVM_HD hvm_pair hvm_new_pair(hvm_port fst, hvm_port snd) { return ((uint64_t)snd << 32) | fst; }
HVM_HD hvm_port hvm_get_fst(hvm_pair p) { return (hvm_port)(p & 0xFFFFFFFFull); }
HVM_HD hvm_port hvm_get_snd(hvm_pair p) { return (hvm_port)(p >> 32); }
HVM_HD int      hvm_is_nod(hvm_port p) { return hvm_get_tag(p) >= HVM_CON; }
HVM_HD int      hvm_is_var(hvm_port p) { return hvm_get_tag(p) == HVM_VAR; }

// `&!` redexes carry bit 28 of the REF value (src/hvm.rs:203-211).
HVM_HD hvm_pair hvm_set_par_flag(hvm_pair p) {
  hvm_port a = hvm_get_fst(p);
  if (hvm_get_tag(a) == HVM_REF) {
    return hvm_new_pair(hvm_new_port(HVM_REF, hvm_get_val(a) | 0x10000000u), hvm_get_snd(p));
  }
  return p;
}

// Rule for an active pair, by tag (src/hvm.c:319-333).  Packed as 8 rows of
// 8 nibbles: row `ta` is a u32 holding the rule for each `tb` at bits 4*tb.
HVM_HD hvm_rule hvm_get_rule(hvm_port a, hvm_port b) {
  // VAR REF ERA NUM CON DUP OPR SWI  (columns, low nibble first)
  const uint32_t rows[8] = {
    0x00000000u, // VAR: LINK everywhere
    0x11112220u, // REF: LINK VOID VOID VOID CALL CALL CALL CALL
    0x33332220u, // ERA: LINK VOID VOID VOID ERAS ERAS ERAS ERAS
    0x76332220u, // NUM: LINK VOID VOID VOID ERAS ERAS OPER SWIT
    0x55543310u, // CON: LINK CALL ERAS ERAS ANNI COMM COMM COMM
    0x55453310u, // DUP: LINK CALL ERAS ERAS COMM ANNI COMM COMM
    0x54556310u, // OPR: LINK CALL ERAS OPER COMM COMM ANNI COMM
    0x45557310u, // SWI: LINK CALL ERAS SWIT COMM COMM COMM ANNI
  };
  return (hvm_rule)((rows[hvm_get_tag(a)] >> (4u * hvm_get_tag(b))) & 0xFu);
}

HVM_HD int hvm_should_swap(hvm_port a, hvm_port b) { return hvm_get_tag(b) < hvm_get_tag(a); }

// LINK, VOID, ERAS, ANNI are high priority (src/hvm.c:345-348).
HVM_HD int hvm_is_high_priority(hvm_rule r) { return (0x1Du >> r) & 1u; }

// ---------------------------------------------------------------------------
// Numb
// ---------------------------------------------------------------------------

HVM_HD hvm_numb hvm_new_sym(uint32_t v) { return (v << 5) | HVM_TY_SYM; }
HVM_HD uint32_t hvm_get_sym(hvm_numb w) { return w >> 5; }
HVM_HD hvm_numb hvm_new_u24(uint32_t v) { return (v << 5) | HVM_TY_U24; }
HVM_HD uint32_t hvm_get_u24(hvm_numb w) { return w >> 5; }
HVM_HD hvm_numb hvm_new_i24(int32_t v) { return ((uint32_t)v << 5) | HVM_TY_I24; }
HVM_HD int32_t  hvm_get_i24(hvm_numb w) { return ((int32_t)(w << 3)) >> 8; }
HVM_HD hvm_tag  hvm_get_typ(hvm_numb w) { return (hvm_tag)(w & 0x1Fu); }

HVM_HD uint32_t hvm_f32_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
HVM_HD float    hvm_bits_f32(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

// f24 packing as the runtimes do it (src/hvm.c:396-406, src/hvm.cu:502-512):
// keep the top 24 bits of the f32 and keep NaNs NaN.  The rounding term uses a
// *logical* not of the kept bits, so an exact tie rounds up unless the kept
// bits are all zero.  Every runtime-produced f24 (OPER, casts) goes through
// this function, so results match the C runtime bit for bit.
HVM_HD hvm_numb hvm_new_f24(float val) {
  uint32_t bits = hvm_f32_bits(val);
  uint32_t shifted = bits >> 8;
  uint32_t lost = bits & 0xFFu;
  uint32_t nan = (val != val) ? 1u : 0u;
  shifted += (nan ^ 1u) & ((lost - ((lost >> 7) & (shifted == 0u ? 1u : 0u))) >> 7);
  shifted |= nan;
  return (shifted << 5) | HVM_TY_F24;
}

// f24 packing as the front end does it for literals (src/hvm.rs:256-265): the
// rounding term uses a *bitwise* not, which is true round-ties-to-even.
HVM_HD hvm_numb hvm_new_f24_literal(float val) {
  uint32_t bits = hvm_f32_bits(val);
  uint32_t shifted = bits >> 8;
  uint32_t lost = bits & 0xFFu;
  uint32_t nan = (val != val) ? 1u : 0u;
  shifted += (nan ^ 1u) & ((lost - ((lost >> 7) & ~shifted)) >> 7);
  shifted |= nan;
  return (shifted << 5) | HVM_TY_F24;
}

HVM_HD float hvm_get_f24(hvm_numb w) { return hvm_bits_f32((w << 3) & 0xFFFFFF00u); }

HVM_HD int hvm_numb_is_num(hvm_numb w) { return hvm_get_typ(w) >= HVM_TY_U24 && hvm_get_typ(w) <= HVM_TY_F24; }
HVM_HD int hvm_numb_is_cast(hvm_numb w) {
  return hvm_get_typ(w) == HVM_TY_SYM && hvm_get_sym(w) >= HVM_TY_U24 && hvm_get_sym(w) <= HVM_TY_F24;
}

// Partial application of an operator symbol (src/hvm.c:426-428).
HVM_HD hvm_numb hvm_partial(hvm_numb a, hvm_numb b) { return (b & ~0x1Fu) | hvm_get_sym(a); }

HVM_HD float hvm_clampf(float x, float lo, float hi) {
  float t = x < lo ? lo : x;
  return t > hi ? hi : t;
}

// Numeric casts (src/hvm.c:437-472).
HVM_HD hvm_numb hvm_cast(hvm_numb a, hvm_numb b) {
  const float U24_MAX = (float)(1 << 24) - 1;
  const float I24_MAX = (float)(1 << 23) - 1;
  const float I24_MIN = (float)(int32_t)(0xFFFFFFFFu << 23);
  uint32_t to = hvm_get_sym(a);
  hvm_tag from = hvm_get_typ(b);
  if (to == HVM_TY_U24) {
    if (from == HVM_TY_U24) return b;
    if (from == HVM_TY_I24) return hvm_new_u24((uint32_t)hvm_get_i24(b));
    if (from == HVM_TY_F24) {
      float v = hvm_get_f24(b);
      if (v != v) return hvm_new_u24(0);
      return hvm_new_u24((uint32_t)hvm_clampf(v, 0.0f, U24_MAX));
    }
  }
  if (to == HVM_TY_I24) {
    if (from == HVM_TY_U24) return hvm_new_i24((int32_t)hvm_get_u24(b));
    if (from == HVM_TY_I24) return b;
    if (from == HVM_TY_F24) {
      float v = hvm_get_f24(b);
      if (v != v) return hvm_new_i24(0);
      return hvm_new_i24((int32_t)hvm_clampf(v, I24_MIN, I24_MAX));
    }
  }
  if (to == HVM_TY_F24) {
    if (from == HVM_TY_U24) return hvm_new_f24((float)hvm_get_u24(b));
    if (from == HVM_TY_I24) return hvm_new_f24((float)hvm_get_i24(b));
    if (from == HVM_TY_F24) return b;
  }
  return hvm_new_u24(0);
}

// The f24 operators of hvm_operate (src/hvm.c:559-579).
HVM_HD_COLD hvm_numb hvm_operate_f24(hvm_tag op, hvm_numb a, hvm_numb b) {
  float av = hvm_get_f24(a), bv = hvm_get_f24(b);
  switch (op) {
    case HVM_OP_ADD: return hvm_new_f24(av + bv);
    case HVM_OP_SUB: return hvm_new_f24(av - bv);
    case HVM_FP_SUB: return hvm_new_f24(bv - av);
    case HVM_OP_MUL: return hvm_new_f24(av * bv);
    case HVM_OP_DIV: return hvm_new_f24(av / bv);
    case HVM_FP_DIV: return hvm_new_f24(bv / av);
    case HVM_OP_REM: return hvm_new_f24(fmodf(av, bv));
    case HVM_FP_REM: return hvm_new_f24(fmodf(bv, av));
    case HVM_OP_EQ:  return hvm_new_u24(av == bv);
    case HVM_OP_NEQ: return hvm_new_u24(av != bv);
    case HVM_OP_LT:  return hvm_new_u24(av < bv);
    case HVM_OP_GT:  return hvm_new_u24(av > bv);
    case HVM_OP_AND: return hvm_new_f24(atan2f(av, bv));
    case HVM_OP_OR:  return hvm_new_f24(logf(bv) / logf(av));
    case HVM_OP_XOR: return hvm_new_f24(powf(av, bv));
    case HVM_OP_SHL: return hvm_new_f24((float)sin((double)(av + bv)));
    case HVM_OP_SHR: return hvm_new_f24((float)tan((double)(av + bv)));
    default:         return hvm_new_f24(0.0f);
  }
}

// The numeric ALU (src/hvm.c:475-582).  u24/i24 division or remainder by zero
// is undefined in the reference C runtime; here it yields 0 instead of trapping.
HVM_HD hvm_numb hvm_operate(hvm_numb a, hvm_numb b) {
  hvm_tag at = hvm_get_typ(a);
  hvm_tag bt = hvm_get_typ(b);
  if (at == HVM_TY_SYM && bt == HVM_TY_SYM) return hvm_new_u24(0);
  if (hvm_numb_is_cast(a) && hvm_numb_is_num(b)) return hvm_cast(a, b);
  if (hvm_numb_is_cast(b) && hvm_numb_is_num(a)) return hvm_cast(b, a);
  if (at == HVM_TY_SYM && bt != HVM_TY_SYM) return hvm_partial(a, b);
  if (at != HVM_TY_SYM && bt == HVM_TY_SYM) return hvm_partial(b, a);
  if (at >= HVM_OP_ADD && bt >= HVM_OP_ADD) return hvm_new_u24(0);
  if (at < HVM_OP_ADD && bt < HVM_OP_ADD) return hvm_new_u24(0);
  hvm_tag op, ty;
  if (at >= HVM_OP_ADD) {
    op = at; ty = bt;
  } else {
    op = bt; ty = at;
    hvm_numb t = a; a = b; b = t;
  }
  if (ty == HVM_TY_U24) {
    uint32_t av = hvm_get_u24(a), bv = hvm_get_u24(b);
    switch (op) {
      case HVM_OP_ADD: return hvm_new_u24(av + bv);
      case HVM_OP_SUB: return hvm_new_u24(av - bv);
      case HVM_FP_SUB: return hvm_new_u24(bv - av);
      case HVM_OP_MUL: return hvm_new_u24(av * bv);
      case HVM_OP_DIV: return hvm_new_u24(bv ? av / bv : 0u);
      case HVM_FP_DIV: return hvm_new_u24(av ? bv / av : 0u);
      case HVM_OP_REM: return hvm_new_u24(bv ? av % bv : 0u);
      case HVM_FP_REM: return hvm_new_u24(av ? bv % av : 0u);
      case HVM_OP_EQ:  return hvm_new_u24(av == bv);
      case HVM_OP_NEQ: return hvm_new_u24(av != bv);
      case HVM_OP_LT:  return hvm_new_u24(av < bv);
      case HVM_OP_GT:  return hvm_new_u24(av > bv);
      case HVM_OP_AND: return hvm_new_u24(av & bv);
      case HVM_OP_OR:  return hvm_new_u24(av | bv);
      case HVM_OP_XOR: return hvm_new_u24(av ^ bv);
      case HVM_OP_SHL: return hvm_new_u24(av << (bv & 31));
      case HVM_FP_SHL: return hvm_new_u24(bv << (av & 31));
      case HVM_OP_SHR: return hvm_new_u24(av >> (bv & 31));
      case HVM_FP_SHR: return hvm_new_u24(bv >> (av & 31));
      default:         return hvm_new_u24(0);
    }
  }
  if (ty == HVM_TY_I24) {
    int32_t av = hvm_get_i24(a), bv = hvm_get_i24(b);
    switch (op) {
      case HVM_OP_ADD: return hvm_new_i24((int32_t)((uint32_t)av + (uint32_t)bv));
      case HVM_OP_SUB: return hvm_new_i24((int32_t)((uint32_t)av - (uint32_t)bv));
      case HVM_FP_SUB: return hvm_new_i24((int32_t)((uint32_t)bv - (uint32_t)av));
      case HVM_OP_MUL: return hvm_new_i24((int32_t)((uint32_t)av * (uint32_t)bv));
      case HVM_OP_DIV: return hvm_new_i24(bv ? av / bv : 0);
      case HVM_FP_DIV: return hvm_new_i24(av ? bv / av : 0);
      case HVM_OP_REM: return hvm_new_i24(bv ? av % bv : 0);
      case HVM_FP_REM: return hvm_new_i24(av ? bv % av : 0);
      case HVM_OP_EQ:  return hvm_new_u24(av == bv);
      case HVM_OP_NEQ: return hvm_new_u24(av != bv);
      case HVM_OP_LT:  return hvm_new_u24(av < bv);
      case HVM_OP_GT:  return hvm_new_u24(av > bv);
      case HVM_OP_AND: return hvm_new_i24(av & bv);
      case HVM_OP_OR:  return hvm_new_i24(av | bv);
      case HVM_OP_XOR: return hvm_new_i24(av ^ bv);
      default:         return hvm_new_i24(0);
    }
  }
  if (ty == HVM_TY_F24) return hvm_operate_f24(op, a, b);
  return hvm_new_u24(0);
}
