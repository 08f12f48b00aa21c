/* ref_main.c — driver for the compiled reference C runtime (oracle/_ref).
 * TEST INFRASTRUCTURE ONLY.  Reads a book buffer file (Book::to_buffer
 * format, src/hvm.rs:957-1003) and calls the reference entry
 * `hvm_c(u32*)` (src/hvm.c:1743), which prints Result/ITRS/TIME/MIPS to
 * stdout exactly as `hvm run-c` does (src/main.rs:74-83). */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void hvm_c(uint32_t* book_buffer);

int main(int argc, char** argv) {
  if (argc < 2) { fprintf(stderr, "usage: %s <file.book>\n", argv[0]); return 2; }
  FILE* f = fopen(argv[1], "rb");
  if (!f) { perror(argv[1]); return 2; }
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  uint32_t* buf = (uint32_t*)malloc((size_t)n + 4);
  if (!buf || fread(buf, 1, (size_t)n, f) != (size_t)n) { fprintf(stderr, "read failed\n"); return 2; }
  fclose(f);
  hvm_c(buf);
  free(buf);
  return 0;
}
