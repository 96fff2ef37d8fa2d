// Writes golden vectors of the C++ standard library's std::mt19937 (seed 0)
// and uniform_real_distribution<double>(0, 1) as used by FRANK's seeded
// generators (src/util/experiment_setup.cpp:12-23, src/functions.cpp:52-66).
// Build and run:  g++ -O2 -o /tmp/mt19937_golden mt19937_golden.cpp && /tmp/mt19937_golden > mt19937_seed0.json
#include <cstdio>
#include <random>

int main() {
  std::mt19937 g;
  g.seed(0);
  std::printf("{\n  \"seed\": 0,\n  \"words\": [");
  for (int i = 0; i < 16; ++i) std::printf("%s%u", i ? ", " : "", (unsigned)g());
  std::mt19937 g2;
  g2.seed(0);
  std::uniform_real_distribution<double> d(0.0, 1.0);
  std::printf("],\n  \"uniform\": [");
  for (int i = 0; i < 16; ++i) std::printf("%s%.17g", i ? ", " : "", d(g2));
  std::mt19937 g3;  // default seed 5489: the standard fixes the 10000th output
  unsigned v = 0;
  for (int i = 0; i < 10000; ++i) v = g3();
  std::printf("],\n  \"default_seed_10000th\": %u\n}\n", v);
  return 0;
}
