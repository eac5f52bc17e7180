/* C++20 host mirror of TitSolver's WCSPH interface over libtit_hip.
 *
 * The reference plugs its step in through C++26 templates
 * (source/tit/sph/time_integrator.hpp:140-184, fluid_equations.hpp:34-69,
 * particle_array.hpp:138-293).  This header keeps their names and the call
 * sequence of titwcsph (source/titwcsph/wcsph.cpp:35-196), so code written
 * against the reference reads the same, and runs every step on the GPU
 * through the C ABI of include/tit_hip.h:
 *
 *   FluidEquations equations{g, mu, domain, containment, TaitEquationOfState{cs_0, rho_0},
 *                            SixthOrderWendlandKernel{}};
 *   SSPRKIntegrator time_integrator{equations, SSPRKOrder::three};
 *   ParticleArray<2> particles;  auto a = particles.append(ParticleType::fluid); ...
 *   ParticleMesh mesh;
 *   for (...) dt = time_integrator.step(mesh, particles);
 *   time_integrator.sync(particles);   // before reading r, v, rho, p
 *
 * Differences from the reference, each forced by the device: the particle
 * state lives on the device between steps (step() uploads it on its first
 * call; sync() copies it back, upload() replaces it); fluid particles come
 * before fixed ones, as the reference's
 * ParticleArray keeps them.  Errors throw Exception, as TIT_ENSURE does
 * (core/exception.hpp:77-85).
 */
#ifndef TIT_AMD_SPH_HPP
#define TIT_AMD_SPH_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../tit_hip.h"

namespace tit_amd::sph {

/// Error of the device path: the status and the context's message.
class Exception : public std::runtime_error {
public:
  Exception(tit_status status, const std::string& what)
      : std::runtime_error(what), status_{status} {}
  auto status() const noexcept -> tit_status { return status_; }

private:
  tit_status status_;
};

template<std::size_t Dim>
using Vec = std::array<double, Dim>;

/// ParticleType (particle_array.hpp:31-35).
enum class ParticleType : std::uint8_t { fluid, fixed };

/// Smoothing kernels (kernel.hpp:444-451).
struct QuarticWendlandKernel {
  static constexpr int id = TIT_KERNEL_WENDLAND_C2;
};
struct SixthOrderWendlandKernel {
  static constexpr int id = TIT_KERNEL_WENDLAND_C4;
};

/// TaitEquationOfState (equation_of_state.hpp:19-72): cs_0, rho_0, gamma.
struct TaitEquationOfState {
  double cs_0;
  double rho_0;
  double gamma = 7.0;
};

/// geom::Surface<Vec<double, 2>> (geom/surface.hpp:30-110): vertices and
/// faces as vertex pairs; a face's normal normalize(cross(b - a)),
/// cross(v) = (v1, -v0), points into the fluid.
struct Surface {
  std::vector<Vec<2>> verts;
  std::vector<std::array<std::uint32_t, 2>> faces;

  void append_vert(const Vec<2>& v) { verts.push_back(v); }
  void append_face(const std::array<std::uint32_t, 2>& f) {
    if (f[0] >= verts.size() || f[1] >= verts.size()) {
      throw Exception(TIT_E_ARG, "Face contains invalid vertex index!");
    }
    faces.push_back(f);
  }
  auto num_verts() const noexcept -> std::size_t { return verts.size(); }
  auto vert(std::size_t i) const -> const Vec<2>& { return verts.at(i); }
};

/// geom::tessellate (geom/tessellation.hpp:28-62): every face split into
/// max(1, ceil(length / d_max)) equal segments, new vertices appended.
inline auto tessellate(const Surface& surf, double d_max) -> Surface {
  if (!(d_max > 0.0)) throw Exception(TIT_E_ARG, "Maximum segment length must be positive.");
  Surface result;
  result.verts = surf.verts;
  for (const auto& [ia, ib] : surf.faces) {
    const Vec<2> a = surf.verts[ia], b = surf.verts[ib];
    const Vec<2> ba{b[0] - a[0], b[1] - a[1]};
    const double d = std::sqrt(ba[1] * ba[1] + ba[0] * ba[0]);
    const auto n = std::max<std::size_t>(1, static_cast<std::size_t>(std::ceil(d / d_max)));
    std::uint32_t prev = ia;
    for (std::size_t i = 1; i < n; ++i) {
      const double t = static_cast<double>(i) / static_cast<double>(n);
      result.verts.push_back({a[0] + t * ba[0], a[1] + t * ba[1]});
      const auto v = static_cast<std::uint32_t>(result.verts.size() - 1);
      result.faces.push_back({prev, v});
      prev = v;
    }
    result.faces.push_back({prev, ib});
  }
  return result;
}

/// Term set of the right-hand sides: the reference's (Ferrari diffusion,
/// physical viscosity mu) or the benchmark's (delta-SPH, Monaghan alpha).
enum class Terms { reference = TIT_FLAVOUR_REF, benchmark = TIT_FLAVOUR_CFG };

/// FluidEquations (fluid_equations.hpp:34-69): constants, optional wall
/// boundary (semi-analytic, 2D: the tessellated domain and the containment
/// surface), equation of state and kernel; the time-step constants of the
/// reference (fluid_equations.hpp:518-520).  Without a domain, the fixed
/// particles are layers of wall particles and need wall normals.
template<class Kernel = QuarticWendlandKernel>
struct FluidEquations {
  double g;
  double mu;
  std::optional<Surface> domain;
  std::optional<Surface> containment;
  TaitEquationOfState eos;
  Kernel kernel{};
  Terms terms = Terms::reference;
  double alpha = 0.0;  // benchmark terms: artificial viscosity
  double delta = 0.0;  // benchmark terms: delta-SPH coefficient
  double cfl = 0.4;
  double c_force = 0.25;
  double c_visc = 0.125;
  /// End every step with post_integrate (shifting and free-surface
  /// correction, fluid_equations.hpp:313-512), as the reference always does
  /// (time_integrator.hpp:181); false leaves it out.
  bool post_integrate = true;
};

/// ParticleArray (particle_array.hpp:138-293) in the reference layout:
/// fluid particles first, then fixed.  Uniform h; per-particle m; r, v, rho,
/// p; wall normals of fixed particles when no wall domain is given.
template<std::size_t Dim>
class ParticleArray {
public:
  static_assert(Dim == 2 || Dim == 3);

  /// Appends a particle and returns its index.
  auto append(ParticleType type) -> std::size_t {
    if (type == ParticleType::fluid && n_fixed_ > 0) {
      throw Exception(TIT_E_ARG, "fluid particles must be appended before fixed ones");
    }
    (type == ParticleType::fluid ? n_fluid_ : n_fixed_) += 1;
    r.push_back({});
    v.push_back({});
    rho.push_back(0.0);
    m.push_back(0.0);
    p.push_back(0.0);
    if (type == ParticleType::fixed) normal.push_back({});
    return r.size() - 1;
  }
  auto size() const noexcept -> std::size_t { return r.size(); }
  auto num_fluid() const noexcept -> std::size_t { return n_fluid_; }
  auto num_fixed() const noexcept -> std::size_t { return n_fixed_; }
  auto has_type(std::size_t a, ParticleType t) const noexcept -> bool {
    return (a < n_fluid_) == (t == ParticleType::fluid);
  }

  double h = 0.0;
  std::vector<Vec<Dim>> r, v;
  std::vector<double> rho, m, p;
  std::vector<Vec<Dim>> normal;  // fixed particles, without a wall domain

private:
  std::size_t n_fluid_ = 0, n_fixed_ = 0;
};

/// ParticleMesh: the device keeps its own cell index, so the mesh carries
/// nothing; it stays in the signature of step().
struct ParticleMesh {};

/// SSPRKOrder (time_integrator.hpp:134-137).
enum class SSPRKOrder { two = 2, three = 3 };

/// SSPRKIntegrator (time_integrator.hpp:140-184) on one GPU.
template<class Equations>
class SSPRKIntegrator {
public:
  explicit SSPRKIntegrator(Equations equations, SSPRKOrder order = SSPRKOrder::three,
                           int device = 0)
      : eq_{std::move(equations)}, order_{order}, device_{device} {}
  ~SSPRKIntegrator() { tit_hip_destroy(ctx_); }
  SSPRKIntegrator(const SSPRKIntegrator&) = delete;
  auto operator=(const SSPRKIntegrator&) -> SSPRKIntegrator& = delete;

  /// Copies the particles to the device (the first step() does it).
  template<std::size_t Dim>
  void upload(const ParticleArray<Dim>& particles) {
    create_(particles);
    std::vector<double> r(particles.size() * Dim), v(particles.size() * Dim);
    for (std::size_t a = 0; a < particles.size(); ++a) {
      for (std::size_t k = 0; k < Dim; ++k) {
        r[a * Dim + k] = particles.r[a][k];
        v[a * Dim + k] = particles.v[a][k];
      }
    }
    std::vector<double> nrm(particles.normal.size() * Dim);
    for (std::size_t e = 0; e < particles.normal.size(); ++e) {
      for (std::size_t k = 0; k < Dim; ++k) nrm[e * Dim + k] = particles.normal[e][k];
    }
    tit_hip_host_view view{};
    view.r = r.data();
    view.v = v.data();
    view.rho = particles.rho.data();
    view.normal = eq_.domain ? nullptr : nrm.data();
    view.n_fluid = particles.num_fluid();
    view.n_fixed = particles.num_fixed();
    // per-particle masses only when they vary (uniform: params.m)
    const bool varying = std::any_of(particles.m.begin(), particles.m.end(),
                                     [&](double x) { return x != particles.m.front(); });
    view.m = varying ? particles.m.data() : nullptr;
    // the cell grid covers the particles' box grown by four kernel widths;
    // particles that leave it are clamped into its border cells
    tit_hip_tank box{};
    for (std::size_t k = 0; k < Dim; ++k) {
      double lo = std::numeric_limits<double>::max(), hi = -lo;
      for (const auto& x : particles.r) {
        lo = std::min(lo, x[k]);
        hi = std::max(hi, x[k]);
      }
      box.lo[k] = lo - 4.0 * particles.h;
      box.hi[k] = hi + 4.0 * particles.h;
    }
    check_(tit_hip_upload(ctx_, &view, &box));
    uploaded_ = true;
  }

  /// SSPRKIntegrator::step: one step on the device; returns dt.
  template<std::size_t Dim>
  auto step(ParticleMesh& /*mesh*/, ParticleArray<Dim>& particles) -> double {
    if (!uploaded_) upload(particles);
    double dt = 0.0;
    check_(tit_hip_step(ctx_, 1, &dt));
    return dt;
  }

  /// Copies r, v, rho, p back (before reading or writing the particles).
  template<std::size_t Dim>
  void sync(ParticleArray<Dim>& particles) {
    if (!uploaded_) throw Exception(TIT_E_STATE, "sync before the first step");
    const std::size_t n = particles.size();
    std::vector<double> r(n * Dim), v(n * Dim);
    check_(tit_hip_download(ctx_, r.data(), v.data(), particles.rho.data(),
                            particles.p.data()));
    for (std::size_t a = 0; a < n; ++a) {
      for (std::size_t k = 0; k < Dim; ++k) {
        particles.r[a][k] = r[a * Dim + k];
        particles.v[a][k] = v[a * Dim + k];
      }
    }
  }

private:
  template<std::size_t Dim>
  void create_(const ParticleArray<Dim>& particles) {
    if (particles.num_fluid() == 0) throw Exception(TIT_E_ARG, "no fluid particles");
    const double m0 = particles.m.front();
    if (ctx_) {
      tit_hip_destroy(ctx_);
      ctx_ = nullptr;
    }
    tit_hip_params p{};
    p.dim = static_cast<int>(Dim);
    p.kernel = decltype(eq_.kernel)::id;
    p.flavour = static_cast<int>(eq_.terms);
    p.ssprk_order = static_cast<int>(order_);
    p.h = particles.h;
    p.m = m0;
    p.rho0 = eq_.eos.rho_0;
    p.cs0 = eq_.eos.cs_0;
    p.xi = eq_.eos.gamma;
    p.g = eq_.g;
    p.mu = eq_.mu;
    p.alpha = eq_.alpha;
    p.delta = eq_.delta;
    p.cfl = eq_.cfl;
    p.c_force = eq_.c_force;
    p.c_visc = eq_.c_visc;
    p.post_integrate = eq_.post_integrate ? 1 : 0;
    check_(tit_hip_create(device_, &p, &ctx_));
    if (eq_.domain) {
      const Surface& d = *eq_.domain;
      const Surface& c = eq_.containment ? *eq_.containment : d;
      std::vector<double> v, cv;
      std::vector<std::uint32_t> f, cf;
      for (const auto& x : d.verts) v.insert(v.end(), {x[0], x[1]});
      for (const auto& e : d.faces) f.insert(f.end(), {e[0], e[1]});
      for (const auto& x : c.verts) cv.insert(cv.end(), {x[0], x[1]});
      for (const auto& e : c.faces) cf.insert(cf.end(), {e[0], e[1]});
      tit_hip_walls w{};
      w.verts = v.data();
      w.n_verts = static_cast<std::uint32_t>(d.verts.size());
      w.n_faces = static_cast<std::uint32_t>(d.faces.size());
      w.faces = f.data();
      w.cverts = cv.data();
      w.n_cverts = static_cast<std::uint32_t>(c.verts.size());
      w.n_cfaces = static_cast<std::uint32_t>(c.faces.size());
      w.cfaces = cf.data();
      check_(tit_hip_set_walls(ctx_, &w));
    }
  }

  void check_(tit_status s) const {
    if (s != TIT_OK) {
      const char* msg = ctx_ ? tit_hip_last_error(ctx_) : nullptr;
      throw Exception(s, std::string("tit_hip: ") +
                             (msg && *msg ? msg : "call failed (is a GPU visible?)"));
    }
  }

  Equations eq_;
  SSPRKOrder order_;
  int device_;
  tit_hip_ctx* ctx_ = nullptr;
  bool uploaded_ = false;
};

} // namespace tit_amd::sph

#endif // TIT_AMD_SPH_HPP
