// The dam break of titwcsph (source/titwcsph/wcsph.cpp:35-196) written against
// the C++ host mirror (include/tit_amd/sph.hpp) and run on the GPU.  Used by
// tests/test_cpp_mirror.py, which steps the oracle from the same initial
// state and compares.
//
//   dam_break_mirror selftest
//       host-only checks of the mirror (tessellation, argument errors)
//   dam_break_mirror walls|layers <dr> <steps> <out.bin>
//       walls:  the closed pool with the semi-analytic wall, reference terms,
//               Wendland C4, h = 2 dr, every step ending with post_integrate
//               (titwcsph's own set-up and step);
//       layers: the open tank with three layers of wall particles, benchmark
//               terms, Wendland C2, h = 2 dr, no post-integration.
//       out.bin: int64 dim, n, n_fluid, n_fixed, steps; then float64 h, m,
//                initial r, v, rho, normals (n_fixed x dim), final r, v, rho,
//                and the step sizes.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tit_amd/sph.hpp"

using namespace tit_amd::sph;

namespace {

constexpr double kG = 9.81;

// Hydrostatic pressure of the water column (wcsph.cpp:129-138).
double column_pressure(double x, double y, double L, double H, double rho_0) {
  double p = rho_0 * kG * (H - y);
  const double pi = 3.14159265358979323846;
  for (int k = 1; k < 100; k += 2) {
    const double kp = static_cast<double>(k) * pi;
    p -= 8 * rho_0 * kG * H / (kp * kp) *
         (std::exp(kp * (x - L) / (2 * H)) * std::cos(kp * y / (2 * H)));
  }
  return p;
}

template<class T>
void put(std::FILE* f, const T* p, std::size_t n) {
  if (n && std::fwrite(p, sizeof(T), n, f) != n) throw std::runtime_error("write failed");
}

void dump_state(std::FILE* f, const ParticleArray<2>& pa) {
  for (const auto& x : pa.r) put(f, x.data(), 2);
  for (const auto& x : pa.v) put(f, x.data(), 2);
  put(f, pa.rho.data(), pa.rho.size());
}

int selftest() {
  int fails = 0;
  auto expect = [&](bool ok, const char* what) {
    if (!ok) {
      std::printf("FAIL %s\n", what);
      ++fails;
    }
  };
  // tessellation of a 1.0 x 0.8 box to 0.03: ceil(1/0.03) = 34 and
  // ceil(0.8/0.03) = 27 pieces per side
  Surface box;
  box.append_vert({0.0, 0.8});
  box.append_vert({1.0, 0.8});
  box.append_vert({1.0, 0.0});
  box.append_vert({0.0, 0.0});
  box.append_face({0, 1});
  box.append_face({1, 2});
  box.append_face({2, 3});
  box.append_face({3, 0});
  const Surface t = tessellate(box, 0.03);
  expect(t.faces.size() == 2 * 34 + 2 * 27, "face count");
  expect(t.verts.size() == t.faces.size(), "vertex count of a closed polyline");
  bool thrown = false;
  try {
    (void)tessellate(box, 0.0);
  } catch (const Exception& e) {
    thrown = e.status() == TIT_E_ARG;
  }
  expect(thrown, "tessellate rejects d_max = 0");
  thrown = false;
  try {
    box.append_face({0, 9});
  } catch (const Exception& e) {
    thrown = e.status() == TIT_E_ARG;
  }
  expect(thrown, "append_face rejects a missing vertex");
  ParticleArray<2> pa;
  (void)pa.append(ParticleType::fluid);
  (void)pa.append(ParticleType::fixed);
  thrown = false;
  try {
    (void)pa.append(ParticleType::fluid);
  } catch (const Exception& e) {
    thrown = e.status() == TIT_E_ARG;
  }
  expect(thrown, "fluid after fixed is rejected");
  expect(pa.num_fluid() == 1 && pa.num_fixed() == 1 && pa.has_type(1, ParticleType::fixed),
         "type ranges");
  std::printf("selftest %s\n", fails ? "failed" : "ok");
  return fails ? 1 : 0;
}

template<class Equations>
int run(Equations equations, ParticleArray<2>& particles, int steps, const char* out) {
  std::FILE* f = std::fopen(out, "wb");
  if (!f) return 2;
  const std::int64_t hdr[5] = {2, static_cast<std::int64_t>(particles.size()),
                               static_cast<std::int64_t>(particles.num_fluid()),
                               static_cast<std::int64_t>(particles.num_fixed()), steps};
  put(f, hdr, 5);
  const double hm[2] = {particles.h, particles.m.front()};
  put(f, hm, 2);
  dump_state(f, particles);
  for (const auto& x : particles.normal) put(f, x.data(), 2);

  SSPRKIntegrator time_integrator{equations, SSPRKOrder::three};
  ParticleMesh mesh;
  std::vector<double> dts;
  for (int s = 0; s < steps; ++s) dts.push_back(time_integrator.step(mesh, particles));
  time_integrator.sync(particles);
  dump_state(f, particles);
  put(f, dts.data(), dts.size());
  std::fclose(f);
  std::printf("%d steps, last dt %.6e\n", steps, dts.empty() ? 0.0 : dts.back());
  return 0;
}

// Closed pool of W x P with a water column L x H in its corner, one wall
// particle per vertex of the tessellated walls (wcsph.cpp:55-142).
int walls_case(double dr, int steps, const char* out) {
  const double L = 0.4, H = 0.5, W = 1.0, P = 0.8;
  const double rho_0 = 1000.0, cs_0 = 20.0 * std::sqrt(kG * H), h_0 = 2.0 * dr;
  const double m_0 = rho_0 * dr * dr, mu = 0.001;
  Surface domain;
  domain.append_vert({0.0, P});
  domain.append_vert({W, P});
  domain.append_vert({W, 0.0});
  domain.append_vert({0.0, 0.0});
  domain.append_face({0, 1});
  domain.append_face({1, 2});
  domain.append_face({2, 3});
  domain.append_face({3, 0});
  domain = tessellate(domain, dr);
  Surface domain2;
  domain2.append_vert({0.0, 0.0});
  domain2.append_vert({W, 0.0});
  domain2.append_vert({W, P});
  domain2.append_vert({0.0, P});
  domain2.append_face({0, 1});
  domain2.append_face({1, 2});
  domain2.append_face({2, 3});
  domain2.append_face({3, 0});
  FluidEquations<SixthOrderWendlandKernel> equations{kG, mu, domain, domain2,
                                                      TaitEquationOfState{cs_0, rho_0}};
  ParticleArray<2> particles;
  const int M = static_cast<int>(std::lround(L / dr)), N = static_cast<int>(std::lround(H / dr));
  for (int i = 0; i < M; ++i) {
    for (int j = 0; j < N; ++j) {
      const auto a = particles.append(ParticleType::fluid);
      particles.r[a] = {dr * (i + 1.0), dr * (j + 1.0)};
    }
  }
  for (std::size_t i = 0; i < domain.num_verts(); ++i) {
    const auto a = particles.append(ParticleType::fixed);
    particles.r[a] = domain.vert(i);
  }
  particles.h = h_0;
  for (std::size_t a = 0; a < particles.size(); ++a) {
    particles.m[a] = m_0;
    if (particles.has_type(a, ParticleType::fixed)) {
      particles.rho[a] = rho_0;
      continue;
    }
    const double p_a = column_pressure(particles.r[a][0], particles.r[a][1], L, H, rho_0);
    particles.rho[a] = rho_0 + p_a / (cs_0 * cs_0);
  }
  return run(equations, particles, steps, out);
}

// Open tank with three layers of wall particles, benchmark terms.
int layers_case(double dr, int steps, const char* out) {
  const double L = 0.4, H = 0.5, W = 1.0, P = 0.8;
  const double rho_0 = 1000.0, cs_0 = 10.0 * std::sqrt(2.0 * kG * H), h_0 = 2.0 * dr;
  FluidEquations<QuarticWendlandKernel> equations{kG, 0.0, std::nullopt, std::nullopt,
                                                  TaitEquationOfState{cs_0, rho_0}};
  equations.terms = Terms::benchmark;
  equations.post_integrate = false;
  equations.alpha = 0.02;
  equations.delta = 0.1;
  equations.cfl = 0.25;
  ParticleArray<2> particles;
  const int M = static_cast<int>(std::lround(L / dr)), N = static_cast<int>(std::lround(H / dr));
  const int nx = static_cast<int>(std::lround(W / dr)), ny = static_cast<int>(std::lround(P / dr));
  for (int i = 0; i < M; ++i) {
    for (int j = 0; j < N; ++j) {
      const auto a = particles.append(ParticleType::fluid);
      particles.r[a] = {dr * (i + 1.0), dr * (j + 1.0)};
    }
  }
  auto wall = [&](int i, int j, Vec<2> n) {
    const auto a = particles.append(ParticleType::fixed);
    particles.r[a] = {dr * i, dr * j};
    const double l = std::sqrt(n[0] * n[0] + n[1] * n[1]);
    particles.normal.back() = {n[0] / l, n[1] / l};
  };
  for (int l = 0; l < 3; ++l) {  // bottom, then the two side walls
    for (int i = -2; i <= nx + 2; ++i) {
      wall(i, -l, {(i <= 0 ? 1.0 : 0.0) - (i >= nx ? 1.0 : 0.0), 1.0});
    }
  }
  for (int l = 0; l < 3; ++l) {
    for (int j = 1; j <= ny; ++j) {
      wall(-l, j, {1.0, 0.0});
      wall(nx + l, j, {-1.0, 0.0});
    }
  }
  particles.h = h_0;
  for (std::size_t a = 0; a < particles.size(); ++a) {
    particles.m[a] = rho_0 * dr * dr;
    if (particles.has_type(a, ParticleType::fixed)) {
      particles.rho[a] = rho_0;
      continue;
    }
    const double p_a = column_pressure(particles.r[a][0], particles.r[a][1], L, H, rho_0);
    particles.rho[a] = rho_0 + p_a / (cs_0 * cs_0);
  }
  return run(equations, particles, steps, out);
}

}  // namespace

int main(int argc, char** argv) {
  try {
    if (argc >= 2 && std::strcmp(argv[1], "selftest") == 0) return selftest();
    if (argc != 5) {
      std::fprintf(stderr, "usage: %s selftest | walls|layers <dr> <steps> <out.bin>\n", argv[0]);
      return 2;
    }
    const double dr = std::atof(argv[2]);
    const int steps = std::atoi(argv[3]);
    if (std::strcmp(argv[1], "walls") == 0) return walls_case(dr, steps, argv[4]);
    if (std::strcmp(argv[1], "layers") == 0) return layers_case(dr, steps, argv[4]);
    std::fprintf(stderr, "unknown case %s\n", argv[1]);
    return 2;
  } catch (const Exception& e) {
    std::fprintf(stderr, "Exception (status %d): %s\n", static_cast<int>(e.status()), e.what());
    return 3;
  }
}
