// X-TED_GPU <labels_file> <adjacency_file>: the reference command line
// (X-TED_GPU/main.cu:4-14): evaluates trees 0 and 1 of the files 20 times,
// one blank line after each run.
#include <cstdio>
#include <exception>

#include "../../include/TED_C++.h"

int main(int argc, char* argv[]) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <labels_file> <adjacency_file>\n", argv[0]);
    return 2;
  }
  try {
    for (int i = 0; i < 20; ++i) {
      test(argv[1], argv[2]);
      std::printf("\n");
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
