// Writes the random point sets of TitSolver's neighbour-search test
// (source/tit/geom/search.test.cpp:131-160): std::mt19937 seeded with 123,
// std::uniform_real_distribution<double>(0, 1), 200 points in 2D drawn first,
// then 200 points in 3D from the same engine, components in order.
// Output: one JSON object on stdout with the points as 17-digit decimals.
#include <cstdio>
#include <random>

int main() {
  std::mt19937 engine{123};
  std::uniform_real_distribution<double> dist{0.0, 1.0};
  std::printf("{\"seed\": 123, \"points2d\": [");
  for (int i = 0; i < 200; ++i) {
    const double x = dist(engine);
    const double y = dist(engine);
    std::printf("%s[%.17g, %.17g]", i ? ", " : "", x, y);
  }
  std::printf("], \"points3d\": [");
  for (int i = 0; i < 200; ++i) {
    const double x = dist(engine);
    const double y = dist(engine);
    const double z = dist(engine);
    std::printf("%s[%.17g, %.17g, %.17g]", i ? ", " : "", x, y, z);
  }
  std::printf("]}\n");
  return 0;
}
