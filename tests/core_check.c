/* Prints the value encodings, rule table and f24 packing of hvm_core.h as
 * JSON, for tests/test_core.py.  Compiled and run by that test. */
#include <stdio.h>
#include <string.h>
#include "../lambdavm-rust_amd/csrc/hvm_core.h"

static unsigned f2u(float f) { unsigned u; memcpy(&u, &f, 4); return u; }
static float u2f(unsigned u) { float f; memcpy(&f, &u, 4); return f; }

int main(int argc, char** argv) {
  printf("{\"rules\": [");
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b)
      printf("%s%d", (a || b) ? "," : "", hvm_get_rule(hvm_new_port(a, 5), hvm_new_port(b, 9)));
  printf("], \"high\": [");
  for (int r = 0; r < 8; ++r) printf("%s%d", r ? "," : "", hvm_is_high_priority(r));
  printf("], \"f24_literal\": [");
  /* argv[1..]: f32 bit patterns; print packed-then-unpacked bits */
  for (int i = 1; i < argc; ++i) {
    unsigned in;
    sscanf(argv[i], "%x", &in);
    printf("%s%u", i > 1 ? "," : "", f2u(hvm_get_f24(hvm_new_f24_literal(u2f(in)))));
  }
  printf("], \"f24_runtime\": [");
  for (int i = 1; i < argc; ++i) {
    unsigned in;
    sscanf(argv[i], "%x", &in);
    printf("%s%u", i > 1 ? "," : "", f2u(hvm_get_f24(hvm_new_f24(u2f(in)))));
  }
  printf("], \"ops\": {");
  /* u24/i24 ALU on the operands used by tests/programs/numerics */
  unsigned ops[] = {HVM_OP_ADD, HVM_OP_SUB, HVM_OP_MUL, HVM_OP_DIV, HVM_OP_REM, HVM_OP_EQ, HVM_OP_NEQ,
                    HVM_OP_LT, HVM_OP_GT, HVM_OP_AND, HVM_OP_OR, HVM_OP_XOR, HVM_OP_SHL, HVM_OP_SHR,
                    HVM_FP_SUB, HVM_FP_DIV, HVM_FP_REM, HVM_FP_SHL, HVM_FP_SHR};
  int first = 1;
  for (unsigned k = 0; k < sizeof ops / sizeof ops[0]; ++k) {
    unsigned op = ops[k];
    /* (10 op 2) via partial application, as `10 ~ $([op] $(2 n))` evaluates */
    unsigned p = hvm_operate(hvm_new_u24(10), hvm_new_sym(op));
    unsigned r = hvm_operate(p, hvm_new_u24(2));
    unsigned pi = hvm_operate(hvm_new_i24(-10), hvm_new_sym(op));
    unsigned ri = hvm_operate(pi, hvm_new_i24(3));
    printf("%s\"%u\": [%u, %u]", first ? "" : ",", op, r, ri);
    first = 0;
  }
  printf("}}\n");
  return 0;
}
