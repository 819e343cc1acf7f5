// Evaluates the generators' elementary functions (aster_amd/csrc/icmath.hpp)
// on the doubles of a binary file, for tests/test_ic.py:
//   icmath_probe <exp|log|sin|cos|pow> <in.bin> <out.bin>
// pow reads (x, y) pairs.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../aster_amd/csrc/icmath.hpp"

int main(int argc, char** argv) {
  if (argc != 4) return 2;
  FILE* f = std::fopen(argv[2], "rb");
  if (!f) return 3;
  std::vector<double> in;
  double buf[4096];
  size_t got;
  while ((got = std::fread(buf, sizeof(double), 4096, f)) > 0) in.insert(in.end(), buf, buf + got);
  std::fclose(f);
  std::vector<double> out;
  const char* fn = argv[1];
  if (!std::strcmp(fn, "pow")) {
    for (size_t i = 0; i + 1 < in.size(); i += 2) out.push_back(aster::icm::pow(in[i], in[i + 1]));
  } else {
    for (double x : in) {
      if (!std::strcmp(fn, "exp")) out.push_back(aster::icm::exp(x));
      else if (!std::strcmp(fn, "log")) out.push_back(aster::icm::log(x));
      else if (!std::strcmp(fn, "sin")) out.push_back(aster::icm::sin(x));
      else if (!std::strcmp(fn, "cos")) out.push_back(aster::icm::cos(x));
      else return 4;
    }
  }
  FILE* g = std::fopen(argv[3], "wb");
  if (!g) return 5;
  std::fwrite(out.data(), sizeof(double), out.size(), g);
  std::fclose(g);
  return 0;
}
