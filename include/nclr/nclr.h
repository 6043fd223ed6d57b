// nclr.h - drop-in C++ API of the MI355X MLS-MPM engine.
//
// Same names, constructor, defaults and accessors as the reference's header-only engine
// (src/nclr.h:20-87): Particle<dim>, Cell<dim>, MaterialModel, MPMSimulation<dim> with
// advance(), particles(), grid(), mu_0, lambda_0 and the k* constants.  The step runs on the
// GPU through the C ABI of include/nclr_mpm.h (link with libnclr_mpm.so); this header holds only
// host-side bookkeeping.
//
// Semantics that differ from the reference, all deliberate:
//  * advance() queues the step on the GPU and returns; particles() and grid() synchronise and
//    copy the state back (cached until the next advance()).  References returned by them stay
//    valid across steps but are refreshed only by the next call.
//  * A particle whose stencil leaves the grid throws std::out_of_range from the next
//    particles()/grid()/synchronize() call (the reference throws from std::vector::at inside
//    advance()); it is raised for every axis, where the reference silently aliases into the
//    neighbouring row on the faster axes.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../nclr_mpm.h"
#include "nclr_math.h"

namespace nclr {

template <int dim>
struct Particle {
    Vector<real, dim> x;  // position
    Vector<real, dim> v;  // velocity
    Matrix<real, dim> F;  // deformation gradient
    Matrix<real, dim> C;  // APIC affine momentum
    real Jp;              // plastic volume ratio
    real mass;
    real volume;
    int c;  // colour (host only)

    Particle(Vector<real, dim> x, int c, Vector<real, dim> v = constvec<dim>(0), real mass = 1.0, real volume = 1.0)
        : x(x), v(v), F(diag<dim>(1)), C(constmat<dim>(0)), Jp(1.0), mass(mass), volume(volume), c(c) {}
};

template <int dim>
struct Cell {
    Vector<real, dim> velocity;
    real mass;
    Cell() : velocity(constvec<dim>(0)), mass(0.0) {}
};

enum class MaterialModel {
    kSnow = 0,
    kJelly,
    kLiquid,
};

namespace detail {
inline void check(int rc) {
    if (rc == MPM_OK) return;
    const std::string msg = mpm_last_error();
    if (rc == MPM_ERR_OUT_OF_RANGE) throw std::out_of_range(msg);
    if (rc == MPM_ERR_INVALID) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}
}  // namespace detail

template <int dim>
class MPMSimulation {
    static_assert(dim == 2 || dim == 3, "MPMSimulation<2> or MPMSimulation<3>");

  public:
    constexpr static int kBoundary = 3;
    constexpr static nclr::real kSnowHardening = 10.0;
    constexpr static nclr::real kJellyHardening = 0.3;
    constexpr static nclr::real kLiquidHardening = 1.0;

    const real mu_0;
    const real lambda_0;

    MPMSimulation(std::vector<Particle<dim>> particles, const MaterialModel model, int res = 64, real dt = 1e-4,
                  real E = 1e4, real nu = 0.2, real gravity = -100, const mpm_options* options = nullptr)
        : mu_0(E / (2 * (1 + nu))),
          lambda_0(E * nu / ((1 + nu) * (1 - 2 * nu))),
          particles_(std::move(particles)),
          res_(res) {
        const size_t n = particles_.size();
        std::vector<float> x(n * dim), v(n * dim), F(n * dim * dim), C(n * dim * dim), Jp(n), m(n), vol(n);
        for (size_t i = 0; i < n; ++i) {
            const auto& p = particles_[i];
            for (int a = 0; a < dim; ++a) {
                x[i * dim + a] = p.x(a);
                v[i * dim + a] = p.v(a);
                for (int b = 0; b < dim; ++b) {
                    F[(i * dim + a) * dim + b] = p.F(a, b);
                    C[(i * dim + a) * dim + b] = p.C(a, b);
                }
            }
            Jp[i] = p.Jp;
            m[i] = p.mass;
            vol[i] = p.volume;
        }
        mpm_handle h = nullptr;
        detail::check(mpm_create(dim, int(model), res, dt, E, nu, gravity, int64_t(n), x.data(), v.data(), F.data(),
                                 C.data(), Jp.data(), m.data(), vol.data(), options, &h));
        handle_.reset(h);
    }

    // advance() (src/nclr.h:80-84): one step, queued on the GPU.
    auto advance() -> void { advance(1); }
    auto advance(int steps) -> void {
        detail::check(mpm_advance(handle_.get(), steps));
        particles_fresh_ = grid_fresh_ = false;
    }

    // particles() (src/nclr.h:86)
    auto particles() const -> const std::vector<Particle<dim>>& {
        if (!particles_fresh_) refresh_particles();
        return particles_;
    }

    // grid() (src/nclr.h:87): the grid after the last step's update ((res+1)^dim cells, flat
    // x-major); empty before the first step, as in the reference.
    auto grid() const -> const std::vector<Cell<dim>>& {
        if (!grid_fresh_) refresh_grid();
        return cells_;
    }

    auto synchronize() const -> void { detail::check(mpm_synchronize(handle_.get())); }
    auto handle() const -> mpm_handle { return handle_.get(); }

  private:
    struct Destroy {
        void operator()(mpm_handle h) const { mpm_destroy(h); }
    };

    void refresh_particles() const {
        const size_t n = particles_.size();
        std::vector<float> x(n * dim), v(n * dim), F(n * dim * dim), C(n * dim * dim), Jp(n);
        detail::check(
            mpm_get_particles(handle_.get(), x.data(), v.data(), F.data(), C.data(), Jp.data(), nullptr, nullptr));
        for (size_t i = 0; i < n; ++i) {
            auto& p = particles_[i];
            for (int a = 0; a < dim; ++a) {
                p.x(a) = x[i * dim + a];
                p.v(a) = v[i * dim + a];
                for (int b = 0; b < dim; ++b) {
                    p.F(a, b) = F[(i * dim + a) * dim + b];
                    p.C(a, b) = C[(i * dim + a) * dim + b];
                }
            }
            p.Jp = Jp[i];
        }
        particles_fresh_ = true;
    }

    void refresh_grid() const {
        if (mpm_steps_taken(handle_.get()) == 0) {
            cells_.clear();
            grid_fresh_ = true;
            return;
        }
        size_t nodes = 1;
        for (int a = 0; a < dim; ++a) nodes *= size_t(res_ + 1);
        std::vector<float> vel(nodes * dim), mass(nodes);
        detail::check(mpm_get_grid(handle_.get(), vel.data(), mass.data()));
        cells_.assign(nodes, Cell<dim>());
        for (size_t i = 0; i < nodes; ++i) {
            for (int a = 0; a < dim; ++a) cells_[i].velocity(a) = vel[i * dim + a];
            cells_[i].mass = mass[i];
        }
        grid_fresh_ = true;
    }

    mutable std::vector<Particle<dim>> particles_;
    mutable std::vector<Cell<dim>> cells_;
    mutable bool particles_fresh_ = true;
    mutable bool grid_fresh_ = true;
    int res_;
    std::unique_ptr<nclr_mpm_sim, Destroy> handle_;
};

}  // namespace nclr
