/*
 * CPU baseline of the WCSPH step -- TEST AND BENCHMARK INFRASTRUCTURE ONLY.
 *
 * The oracle's SSPRK step (tit_oracle.c) laid out the way the reference runs
 * it on a multicore host: every loop over particles is a parallel loop over
 * a pool of std::threads standing in for TBB's par::for_each
 * (source/tit/par/algorithms.hpp:57-177), and the pair loops scatter over
 * unique pairs b < a from colour-separated blocks, as par::block_for_each
 * over ParticleMesh::block_pairs does (particle_mesh.hpp:164-241,
 * algorithms.hpp:84-97).  The blocks here are x-slabs at least one support
 * radius wide: pairs inside a slab run in parallel over slabs, then the
 * pairs across the even slab interfaces, then those across the odd ones, so
 * no two threads write one particle.  The reference colours a recursive
 * inertial bisection instead; only the summation order differs.
 *
 * Scope: the wall-free form of the step (fixed particles carry a wall
 * normal), both term sets, Wendland C2/C4, 2D and 3D, SSPRK order 2 or 3,
 * no post-integration.  Per pair and per particle the arithmetic is the
 * oracle's, so a step agrees with orc_ssprk_step to summation-order
 * rounding; the CPU test suite checks that.
 *
 * Only the benchmark's CPU-baseline leg and the tests load this library.
 */
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "tit_oracle.h"

namespace {

/* ---- thread pool ------------------------------------------------------- */

class Pool {
 public:
  explicit Pool(int n_threads) {
    for (int t = 1; t < n_threads; ++t) workers_.emplace_back([this] { work(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
  }
  int size() const { return static_cast<int>(workers_.size()) + 1; }

  /* fn(i) for every i in [0, n_items); items go out one at a time from an
   * atomic counter; returns when all have run. */
  void run(size_t n_items, const std::function<void(size_t)>& fn) {
    if (n_items == 0) return;
    if (workers_.empty() || n_items == 1) {
      for (size_t i = 0; i < n_items; ++i) fn(i);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      n_items_ = n_items;
      next_.store(0);
      active_ = static_cast<int>(workers_.size());
      ++gen_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return active_ == 0; });
    fn_ = nullptr;
  }

 private:
  void drain() {
    for (size_t i = next_.fetch_add(1); i < n_items_; i = next_.fetch_add(1)) (*fn_)(i);
  }
  void work() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
      }
      drain();
      std::lock_guard<std::mutex> lk(mu_);
      if (--active_ == 0) done_.notify_one();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(size_t)>* fn_ = nullptr;
  size_t n_items_ = 0;
  std::atomic<size_t> next_{0};
  int active_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

/* Splits [0, n) into ranges of `grain` and runs fn(begin, end) on the pool. */
void parallel_for(Pool& pool, size_t n, size_t grain,
                  const std::function<void(size_t, size_t)>& fn) {
  const size_t items = (n + grain - 1) / grain;
  pool.run(items, [&](size_t i) { fn(i * grain, std::min(n, (i + 1) * grain)); });
}

/* ---- kernels and EOS (sph/kernel.hpp:126-163, equation_of_state.hpp) -- */

template <int D>
inline double dot(const double* a, const double* b) {
  double r = a[0] * b[0];
  for (int k = 1; k < D; ++k) r += a[k] * b[k];
  return r;
}

struct Consts {
  int kernel, flavour;
  double h, h_inv, w_norm, m, rho0, cs0, xi, g, mu, alpha, delta, cfl, c_force, c_visc;
  double B;
};

/* 1 / (zero moment * unit sphere area), sph/kernel.hpp:126-131 */
inline double kernel_weight(int kernel, int dim) {
  const double pi = 3.14159265358979323846;
  if (kernel == ORC_KERNEL_WENDLAND_C2) return dim == 2 ? 7.0 / (4.0 * pi) : 21.0 / (16.0 * pi);
  return dim == 2 ? 9.0 / (4.0 * pi) : 495.0 / (256.0 * pi);
}

inline double unit_deriv(int kernel, double q) {
  if (!(q < 2.0)) return 0.0;
  const double t = 1.0 - q / 2.0;
  if (kernel == ORC_KERNEL_WENDLAND_C2) return -5.0 * q * (t * t * t);
  const double t5 = t * t * t * t * t;
  return -14.0 / 3.0 * q * (1.0 + 2.5 * q) * t5;
}

inline double unit_value(int kernel, double q) {
  if (!(q < 2.0)) return 0.0;
  const double t = 1.0 - q / 2.0;
  if (kernel == ORC_KERNEL_WENDLAND_C2) return (1.0 + 2.0 * q) * (t * t) * (t * t);
  const double t3 = t * t * t;
  return (1.0 + 3.0 * q + 35.0 / 12.0 * q * q) * (t3 * t3);
}

template <int D>
inline void kernel_grad(const Consts& c, const double* x, double* out) {
  const double s2 = dot<D>(x, x);
  const double rn = std::sqrt(s2);
  const double q = c.h_inv * rn;
  const double tiny = std::cbrt(DBL_EPSILON);
  const double s = c.w_norm * unit_deriv(c.kernel, q);
  if (s2 < tiny * tiny) { /* normalize(0) = 0 (core/_vec/vec.hpp:663-679) */
    for (int k = 0; k < D; ++k) out[k] = 0.0;
    return;
  }
  for (int k = 0; k < D; ++k) out[k] = s * ((x[k] / rn) * c.h_inv);
}

template <int D>
inline double kernel_value(const Consts& c, const double* x) {
  return c.w_norm * unit_value(c.kernel, c.h_inv * std::sqrt(dot<D>(x, x)));
}

inline double pressure(const Consts& c, double rho) {
  return c.B * (std::pow(rho / c.rho0, c.xi) - 1.0);
}
inline double sound_speed(const Consts& c, double rho) {
  return c.cs0 * std::pow(rho / c.rho0, (c.xi - 1.0) / 2.0);
}
inline double potential(const Consts& c, double rho) {
  const double xi_1 = c.xi - 1.0;
  return c.cs0 * c.cs0 * (std::pow(rho / c.rho0, xi_1) - 1.0) / xi_1;
}
inline double density_from_potential(const Consts& c, double H) {
  const double xi_1 = c.xi - 1.0;
  return c.rho0 * std::pow(1.0 + xi_1 * H / (c.cs0 * c.cs0), 1.0 / xi_1);
}

/* ---- the step ----------------------------------------------------------- */

struct Workspace {
  std::unique_ptr<Pool> pool;
  std::vector<uint64_t> off;
  std::vector<uint32_t> idx;
  std::vector<std::vector<uint32_t>> block_idx;
  std::vector<uint64_t> block_len;
  std::vector<size_t> cell_off, cell_fill;
  std::vector<uint32_t> cell_pts, cell_of;
  std::vector<uint32_t> slab_of, slab_pts;
  std::vector<size_t> slab_off;
  std::vector<double> r_old, v_old, rho_old;
};

Workspace g_ws;

constexpr size_t kGrain = 4096;

/* GridIndex over the current positions (grid_search.hpp:73-113) and the
 * sorted adjacency of every particle (ParticleMesh::search_,
 * particle_mesh.hpp:123-162), the per-particle queries run in parallel
 * blocks whose lists are then stitched in particle order. */
template <int D>
void build_adjacency(Workspace& w, size_t n, const double* pts, double radius, double hint) {
  Pool& pool = *w.pool;
  double lo[3], hi[3], cext[3], inv[3];
  size_t nc[3] = {1, 1, 1};
  for (int k = 0; k < D; ++k) lo[k] = hi[k] = pts[k];
  for (size_t i = 1; i < n; ++i) {
    for (int k = 0; k < D; ++k) {
      lo[k] = std::min(lo[k], pts[i * D + k]);
      hi[k] = std::max(hi[k], pts[i * D + k]);
    }
  }
  size_t ncells = 1;
  for (int k = 0; k < D; ++k) {
    lo[k] -= hint / 2.0;
    hi[k] += hint / 2.0;
    const double ext = hi[k] - lo[k];
    const double ncf = std::max(1.0, std::ceil(ext / hint));
    nc[k] = static_cast<size_t>(ncf);
    cext[k] = ext / ncf;
    inv[k] = 1.0 / cext[k];
    ncells *= nc[k];
  }
  auto flat = [&](const size_t* c) {
    size_t f = c[0];
    for (int k = 1; k < D; ++k) f = nc[k] * f + c[k];
    return f;
  };
  w.cell_of.resize(n);
  w.cell_off.assign(ncells + 1, 0);
  parallel_for(pool, n, kGrain, [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) {
      size_t c[3];
      for (int k = 0; k < D; ++k) c[k] = static_cast<size_t>((pts[i * D + k] - lo[k]) * inv[k]);
      w.cell_of[i] = static_cast<uint32_t>(flat(c));
    }
  });
  for (size_t i = 0; i < n; ++i) w.cell_off[w.cell_of[i] + 1] += 1;
  for (size_t c = 0; c < ncells; ++c) w.cell_off[c + 1] += w.cell_off[c];
  w.cell_fill.assign(w.cell_off.begin(), w.cell_off.end() - 1);
  w.cell_pts.resize(n);
  for (size_t i = 0; i < n; ++i) w.cell_pts[w.cell_fill[w.cell_of[i]]++] = static_cast<uint32_t>(i);

  const double r2 = radius * radius;
  const size_t n_blocks = (n + kGrain - 1) / kGrain;
  w.block_idx.resize(n_blocks);
  w.block_len.assign(n_blocks + 1, 0);
  w.off.resize(n + 1);
  pool.run(n_blocks, [&](size_t blk) {
    std::vector<uint32_t>& out = w.block_idx[blk];
    out.clear();
    const size_t a0 = blk * kGrain, a1 = std::min(n, a0 + kGrain);
    for (size_t a = a0; a < a1; ++a) {
      const double* ctr = pts + a * D;
      size_t clo[3] = {0, 0, 0}, chi[3] = {0, 0, 0};
      for (int k = 0; k < D; ++k) {
        const double half = cext[k] / 2.0;
        double l = ctr[k] - radius - half, u = ctr[k] + radius + half;
        l = std::max(l, lo[k]) + half;
        u = std::min(u, hi[k]) - half;
        clo[k] = static_cast<size_t>((l - lo[k]) * inv[k]);
        chi[k] = static_cast<size_t>((u - lo[k]) * inv[k]);
      }
      const size_t start = out.size();
      size_t c[3];
      for (c[0] = clo[0]; c[0] <= chi[0]; ++c[0]) {
        for (c[1] = clo[1]; c[1] <= chi[1]; ++c[1]) {
          const size_t z_lo = D == 3 ? clo[2] : 0, z_hi = D == 3 ? chi[2] : 0;
          for (size_t z = z_lo; z <= z_hi; ++z) {
            c[2] = z;
            const size_t f = flat(c);
            for (size_t j = w.cell_off[f]; j < w.cell_off[f + 1]; ++j) {
              const uint32_t b = w.cell_pts[j];
              double d[3];
              for (int k = 0; k < D; ++k) d[k] = pts[size_t(b) * D + k] - ctr[k];
              if (dot<D>(d, d) <= r2) out.push_back(b);
            }
          }
        }
      }
      std::sort(out.begin() + start, out.end());
      w.off[a + 1] = out.size() - start; /* list length for now */
    }
    w.block_len[blk + 1] = out.size();
  });
  for (size_t b = 0; b < n_blocks; ++b) w.block_len[b + 1] += w.block_len[b];
  w.idx.resize(w.block_len[n_blocks]);
  w.off[0] = 0;
  for (size_t a = 0; a < n; ++a) w.off[a + 1] += w.off[a];
  pool.run(n_blocks, [&](size_t blk) {
    std::copy(w.block_idx[blk].begin(), w.block_idx[blk].end(),
              w.idx.begin() + static_cast<std::ptrdiff_t>(w.block_len[blk]));
  });
}

/* x-slabs of width >= the support radius: particles of slab s, and the slab
 * of every particle. */
template <int D>
void build_slabs(Workspace& w, size_t n, const double* r, double radius) {
  double xlo = r[0], xhi = r[0];
  for (size_t i = 1; i < n; ++i) {
    xlo = std::min(xlo, r[i * D]);
    xhi = std::max(xhi, r[i * D]);
  }
  const size_t n_slabs = std::max<size_t>(1, static_cast<size_t>((xhi - xlo) / radius));
  const double width = (xhi - xlo) / static_cast<double>(n_slabs);
  w.slab_of.resize(n);
  w.slab_off.assign(n_slabs + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    size_t s = width > 0.0 ? static_cast<size_t>((r[i * D] - xlo) / width) : 0;
    if (s >= n_slabs) s = n_slabs - 1;
    w.slab_of[i] = static_cast<uint32_t>(s);
    w.slab_off[s + 1] += 1;
  }
  for (size_t s = 0; s < n_slabs; ++s) w.slab_off[s + 1] += w.slab_off[s];
  std::vector<size_t> fill(w.slab_off.begin(), w.slab_off.end() - 1);
  w.slab_pts.resize(n);
  for (size_t i = 0; i < n; ++i) w.slab_pts[fill[w.slab_of[i]]++] = static_cast<uint32_t>(i);
}

/* Runs pair(a, b) over every unique pair b < a of the adjacency so that no
 * two threads touch one particle: pairs inside a slab (parallel over slabs),
 * then pairs across interfaces 0, 2, 4, ..., then 1, 3, 5, ...  A slab is at
 * least one support radius wide, so neighbours are at most one slab apart. */
template <class F>
void for_each_pair(Workspace& w, const F& pair) {
  const size_t n_slabs = w.slab_off.size() - 1;
  auto visit = [&](size_t s, int mode, size_t iface) {
    for (size_t j = w.slab_off[s]; j < w.slab_off[s + 1]; ++j) {
      const uint32_t a = w.slab_pts[j];
      for (uint64_t q = w.off[a]; q < w.off[a + 1]; ++q) {
        const uint32_t b = w.idx[q];
        if (b >= a) break;
        const uint32_t sb = w.slab_of[b];
        if (mode == 0 ? sb == s : std::min<size_t>(sb, s) == iface && sb != s) pair(a, b);
      }
    }
  };
  w.pool->run(n_slabs, [&](size_t s) { visit(s, 0, 0); });
  for (size_t parity = 0; parity < 2; ++parity) {
    const size_t n_if = n_slabs > parity + 1 ? (n_slabs - 1 - parity + 1) / 2 : 0;
    w.pool->run(n_if, [&](size_t i) {
      const size_t iface = parity + 2 * i;
      visit(iface, 1, iface);
      visit(iface + 1, 1, iface);
    });
  }
}

template <int D>
int ssprk_step(const orc_params* prm, orc_state* s, int order, double* dt_out) {
  Workspace& w = g_ws;
  Pool& pool = *w.pool;
  Consts c{};
  c.kernel = prm->kernel;
  c.flavour = prm->flavour;
  c.h = prm->h;
  c.h_inv = 1.0 / prm->h;
  c.w_norm = kernel_weight(prm->kernel, D);
  for (int k = 0; k < D; ++k) c.w_norm *= c.h_inv;
  c.m = prm->m;
  c.rho0 = prm->rho0;
  c.cs0 = prm->cs0;
  c.xi = prm->xi;
  c.g = prm->g;
  c.mu = prm->mu;
  c.alpha = prm->alpha;
  c.delta = prm->delta;
  c.cfl = prm->cfl;
  c.c_force = prm->c_force;
  c.c_visc = prm->c_visc;
  c.B = prm->rho0 * prm->cs0 * prm->cs0 / prm->xi;
  const size_t nf = s->n_fluid, n = s->n_fluid + s->n_fixed;
  const double radius = 2.0 * prm->h;
  double* r = s->r;
  double* v = s->v;
  double* rho = s->rho;
  w.r_old.assign(r, r + nf * D);
  w.v_old.assign(v, v + nf * D);
  w.rho_old.assign(rho, rho + nf);
  double dt = 0.0;
  for (int stage = 1; stage <= order; ++stage) {
    /* prepare: index + setup_boundary (fluid_equations.hpp:99-164) */
    build_adjacency<D>(w, n, r, radius, prm->h);
    parallel_for(pool, n - nf, 256, [&](size_t b0, size_t b1) {
      for (size_t e = nf + b0; e < nf + b1; ++e) {
        for (int k = 0; k < D; ++k) v[e * D + k] = 0.0;
        const double* nr = s->normal + (e - nf) * D;
        double n_e[3] = {0.0, 0.0, 0.0};
        const double nn = std::sqrt(dot<D>(nr, nr));
        const double tiny = std::cbrt(DBL_EPSILON);
        if (!(dot<D>(nr, nr) < tiny * tiny)) {
          for (int k = 0; k < D; ++k) n_e[k] = nr[k] / nn;
        }
        double S_e = 0.0, H_e = 0.0;
        for (uint64_t j = w.off[e]; j < w.off[e + 1]; ++j) {
          const size_t b = w.idx[j];
          if (b >= nf) continue;
          const double V_b = c.m / rho[b];
          double r_be[3];
          for (int k = 0; k < D; ++k) r_be[k] = r[b * D + k] - r[e * D + k];
          const double W_be = kernel_value<D>(c, r_be);
          S_e += V_b * W_be;
          H_e += V_b * (potential(c, rho[b]) + c.g * dot<D>(r_be, n_e) * n_e[1]) * W_be;
        }
        rho[e] = density_from_potential(c, std::fabs(S_e) <= tiny ? 0.0 : H_e / S_e);
      }
    });
    /* compute_time_step (fluid_equations.hpp:199-222), at stage 1 */
    if (stage == 1) {
      const size_t n_blk = (nf + kGrain - 1) / kGrain;
      std::vector<double> part(std::max<size_t>(n_blk, 1), DBL_MAX);
      pool.run(n_blk, [&](size_t blk) {
        double m = DBL_MAX;
        for (size_t a = blk * kGrain; a < std::min(nf, (blk + 1) * kGrain); ++a) {
          const double vn = std::sqrt(dot<D>(v + a * D, v + a * D));
          m = std::min(m, c.cfl * c.h / (sound_speed(c, rho[a]) + vn));
          if (c.mu > 0.0) m = std::min(m, c.c_visc * c.h * c.h * rho[a] / c.mu);
          double an = std::sqrt(dot<D>(s->dv_dt + a * D, s->dv_dt + a * D));
          if (an < c.g) an = c.g;
          m = std::min(m, c.c_force * std::sqrt(c.h / an));
        }
        part[blk] = m;
      });
      dt = *std::min_element(part.begin(), part.end());
    }
    /* EOS and zeroed derivatives */
    parallel_for(pool, n, kGrain, [&](size_t b0, size_t b1) {
      for (size_t a = b0; a < b1; ++a) {
        s->cs[a] = sound_speed(c, rho[a]);
        s->p[a] = pressure(c, rho[a]);
        s->drho_dt[a] = 0.0;
        for (int k = 0; k < D; ++k) s->dv_dt[a * D + k] = 0.0;
        if (a < nf) s->dv_dt[a * D + 1] = -c.g;
      }
    });
    build_slabs<D>(w, n, r, radius);
    /* compute_continuity (fluid_equations.hpp:232-260) */
    for_each_pair(w, [&](size_t a, size_t b) {
      double r_ab[3], v_ab[3], grad[3];
      for (int k = 0; k < D; ++k) {
        r_ab[k] = r[a * D + k] - r[b * D + k];
        v_ab[k] = v[a * D + k] - v[b * D + k];
      }
      kernel_grad<D>(c, r_ab, grad);
      const double rho_ab = rho[a] - rho[b];
      if (c.flavour == ORC_FLAVOUR_REF) {
        const double cs_ab = std::max(s->cs[a], s->cs[b]);
        const double rn = std::sqrt(dot<D>(r_ab, r_ab));
        double ta[3], tb[3];
        for (int k = 0; k < D; ++k) {
          const double Psi = cs_ab * rho_ab * r_ab[k] / rn;
          ta[k] = v_ab[k] + Psi / rho[b];
          tb[k] = -v_ab[k] + Psi / rho[a];
        }
        s->drho_dt[a] += c.m * dot<D>(ta, grad);
        s->drho_dt[b] -= c.m * dot<D>(tb, grad);
      } else {
        const double r2 = dot<D>(r_ab, r_ab);
        const double coef = 2.0 * rho_ab / r2;
        double psi[3];
        for (int k = 0; k < D; ++k) psi[k] = coef * r_ab[k];
        const double dhc = c.delta * c.h * c.cs0;
        const double vg = dot<D>(v_ab, grad), pg = dot<D>(psi, grad);
        s->drho_dt[a] += c.m * vg + dhc * (c.m / rho[b]) * pg;
        s->drho_dt[b] += c.m * vg - dhc * (c.m / rho[a]) * pg;
      }
    });
    /* compute_momentum (fluid_equations.hpp:267-305) */
    for_each_pair(w, [&](size_t a, size_t b) {
      double r_ab[3], v_ab[3], grad[3];
      for (int k = 0; k < D; ++k) {
        r_ab[k] = r[a * D + k] - r[b * D + k];
        v_ab[k] = v[a * D + k] - v[b * D + k];
      }
      kernel_grad<D>(c, r_ab, grad);
      const double P_ab = s->p[a] / (rho[a] * rho[a]) + s->p[b] / (rho[b] * rho[b]);
      double Pi_ab = 0.0;
      const double vr = dot<D>(v_ab, r_ab);
      if (c.flavour == ORC_FLAVOUR_REF) {
        Pi_ab = 2.0 * c.mu * vr / (rho[a] * rho[b] * dot<D>(r_ab, r_ab));
      } else if (vr < 0.0) {
        const double c_ab = 0.5 * (s->cs[a] + s->cs[b]);
        const double rho_m = 0.5 * (rho[a] + rho[b]);
        Pi_ab = c.alpha * c_ab * c.h * vr / (rho_m * (dot<D>(r_ab, r_ab) + 0.01 * c.h * c.h));
      }
      const double f = c.m * (Pi_ab - P_ab);
      for (int k = 0; k < D; ++k) {
        s->dv_dt[a * D + k] += f * grad[k];
        s->dv_dt[b * D + k] -= f * grad[k];
      }
    });
    /* substep update and Shu-Osher combination (time_integrator.hpp:159-223) */
    double wgt = 0.0;
    if (order == 2 && stage == 2) wgt = 1.0 / 2.0;
    if (order == 3 && stage == 2) wgt = 1.0 / 4.0;
    if (order == 3 && stage == 3) wgt = 2.0 / 3.0;
    parallel_for(pool, nf, kGrain, [&](size_t b0, size_t b1) {
      for (size_t a = b0; a < b1; ++a) {
        for (int k = 0; k < D; ++k) r[a * D + k] += dt * v[a * D + k];
        for (int k = 0; k < D; ++k) v[a * D + k] += dt * s->dv_dt[a * D + k];
        rho[a] += dt * s->drho_dt[a];
        if (wgt != 0.0) {
          for (int k = 0; k < D; ++k) {
            r[a * D + k] = (1.0 - wgt) * w.r_old[a * D + k] + wgt * r[a * D + k];
            v[a * D + k] = (1.0 - wgt) * w.v_old[a * D + k] + wgt * v[a * D + k];
          }
          rho[a] = (1.0 - wgt) * w.rho_old[a] + wgt * rho[a];
        }
      }
    });
  }
  if (dt_out) *dt_out = dt;
  return 0;
}

}  // namespace

extern "C" {

/* Sets the pool to n_threads threads (1: the caller alone).  Returns the
 * thread count in use. */
int cpu_set_threads(int n_threads) {
  if (n_threads < 1) n_threads = 1;
  if (!g_ws.pool || g_ws.pool->size() != n_threads) {
    g_ws.pool.reset();
    g_ws.pool = std::make_unique<Pool>(n_threads);
  }
  return g_ws.pool->size();
}

/* One SSPRK step of the wall-free WCSPH equations on the thread pool;
 * the same contract as orc_ssprk_step.  Returns -1 for walls,
 * post-integration or an unsupported order/dimension. */
int cpu_ssprk_step(const orc_params* prm, orc_state* s, int order, double* dt_out) {
  if (order != 2 && order != 3) return -1;
  if (prm->walls || prm->post_integrate) return -1;
  if (s->n_fluid + s->n_fixed == 0) return -1;
  if (!g_ws.pool) cpu_set_threads(1);
  if (prm->dim == 2) return ssprk_step<2>(prm, s, order, dt_out);
  if (prm->dim == 3) return ssprk_step<3>(prm, s, order, dt_out);
  return -1;
}

}  // extern "C"
