// Substitution-matrix file reader (the blosum64.txt format of the reference:
// 24 rows of 24 comma-separated integers, letter order
// ARNDCQEGHILKMFPSTWYVBZX*, blosum64.txt:1-24). The reference ships the file
// but reads it nowhere; the -b option reads it here.
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <vector>

#include "Runtime.h"

namespace mstrt {

bool loadMatrix(const std::string& path, int8_t out[576]) {
  std::ifstream in(path);
  if (!in) return false;
  std::vector<int> v;
  std::string line;
  while (std::getline(in, line)) {
    std::stringstream ss(line);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
      size_t a = tok.find_first_not_of(" \t\r");
      if (a == std::string::npos) continue;  // empty field (a trailing comma)
      char* end = nullptr;
      const long x = strtol(tok.c_str() + a, &end, 10);
      if (end == tok.c_str() + a) return false;  // not a number
      for (; *end; ++end)
        if (*end != ' ' && *end != '\t' && *end != '\r') return false;
      v.push_back((int)(x < -1000 ? -1000 : (x > 1000 ? 1000 : x)));
    }
  }
  if (v.size() != 576) return false;
  for (int k = 0; k < 576; ++k) {
    if (v[k] < -128 || v[k] > 127) return false;
    out[k] = (int8_t)v[k];
  }
  return true;
}

}  // namespace mstrt
