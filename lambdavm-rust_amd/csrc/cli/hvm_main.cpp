// hvm_main.cpp — command-line host: `hvm run-hip <file.hvm>`.
//
// Mirrors the reference CLI's run-cu path (src/main.rs:87-96): read the file,
// parse and build the book, serialise it with Book::to_buffer, and hand the
// words to the GPU runtime through its C entry (`hvm_hip`, the drop-in for
// `hvm_cu`).  Parse errors and unbound references end the process the way the
// reference's panics do (message on stderr, exit status 101).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../../include/hvm_hip.h"

static int usage() {
  fprintf(stderr,
          "HVM2 on MI355X\n\n"
          "Usage: hvm <COMMAND> <file>\n\n"
          "Commands:\n"
          "  run-hip   Interprets a file (using HIP on the GPU)\n"
          "  compile   Writes the book buffer of a file (Book::to_buffer format) to stdout\n");
  return 2;
}

int main(int argc, char** argv) {
  if (argc < 3) return usage();
  std::string cmd = argv[1];
  std::ifstream f(argv[2], std::ios::binary);
  if (!f) {
    fprintf(stderr, "thread 'main' panicked:\nUnable to read file\n");
    return 101;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  std::string code = ss.str();
  uint32_t* words = nullptr;
  size_t n = 0;
  char* err = nullptr;
  if (hvm_front_compile(code.c_str(), &words, &n, &err) != 0) {
    fprintf(stderr, "thread 'main' panicked:\n%s\n", err ? err : "parse error");
    hvm_hip_free(err);
    return 101;
  }
  if (cmd == "run-hip") {
    hvm_hip(words);
  } else if (cmd == "compile") {
    fwrite(words, 4, n, stdout);
  } else {
    hvm_hip_free(words);
    return usage();
  }
  hvm_hip_free(words);
  return 0;
}
