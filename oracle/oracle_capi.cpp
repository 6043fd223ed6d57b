// Test infrastructure only: C entry points over the CPU oracle
// (nclr_oracle.hpp) for the Python tests and bench.py's cpu_baseline leg.
// Arrays are particle-major: x[n][d], v[n][d], F[n][d][d] (row-major),
// C[n][d][d], Jp[n], mass[n], volume[n].  Grids use the reference's flat
// x-major node order (src/nclr.h:141-142,152): vel[nodes][d], mass[nodes].
#include <chrono>
#include <cstring>
#include <string>

#include "nclr_adjoint.hpp"
#include "nclr_oracle.hpp"

namespace {

thread_local std::string g_err;

template <typename T, int D>
std::vector<oracle::Particle<T, D>> load(int64_t n, const T* x, const T* v, const T* F, const T* C, const T* Jp,
                                         const T* mass, const T* volume) {
    std::vector<oracle::Particle<T, D>> ps(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        auto& p = ps[size_t(i)];
        for (int a = 0; a < D; ++a) {
            p.x[a] = x[i * D + a];
            p.v[a] = v[i * D + a];
            for (int b = 0; b < D; ++b) {
                p.F[a][b] = F[(i * D + a) * D + b];
                p.C[a][b] = C[(i * D + a) * D + b];
            }
        }
        p.Jp = Jp[i];
        p.mass = mass[i];
        p.volume = volume[i];
        p.c = 0;
    }
    return ps;
}

template <typename T, int D>
void store(const std::vector<oracle::Particle<T, D>>& ps, T* x, T* v, T* F, T* C, T* Jp) {
    for (size_t i = 0; i < ps.size(); ++i) {
        const auto& p = ps[i];
        for (int a = 0; a < D; ++a) {
            x[i * D + a] = p.x[a];
            v[i * D + a] = p.v[a];
            for (int b = 0; b < D; ++b) {
                F[(i * D + a) * D + b] = p.F[a][b];
                C[(i * D + a) * D + b] = p.C[a][b];
            }
        }
        Jp[i] = p.Jp;
    }
}

template <typename T, int D>
void store_grid(const std::vector<oracle::Cell<T, D>>& cells, T* vel, T* mass) {
    if (!vel && !mass) return;
    for (size_t i = 0; i < cells.size(); ++i) {
        if (vel)
            for (int a = 0; a < D; ++a) vel[i * D + a] = cells[i].v[a];
        if (mass) mass[i] = cells[i].m;
    }
}

template <typename T, int D>
void load_grid(std::vector<oracle::Cell<T, D>>& cells, const T* vel, const T* mass) {
    for (size_t i = 0; i < cells.size(); ++i) {
        for (int a = 0; a < D; ++a) cells[i].v[a] = vel[i * D + a];
        cells[i].m = mass[i];
    }
}

template <typename Fn>
int guarded(Fn&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::out_of_range& e) {
        g_err = std::string("out_of_range: ") + e.what();
        return 1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 2;
    }
}

template <typename T, int D>
int run(int material, int res, T dt, T E, T nu, T g, int idm, int64_t n, int nsteps, T* x, T* v, T* F, T* C, T* Jp,
        const T* mass, const T* volume, T* grid_v, T* grid_m, int threads = 1) {
    return guarded([&] {
        oracle::Simulation<T, D> sim(load<T, D>(n, x, v, F, C, Jp, mass, volume), material, res, dt, E, nu, g, idm);
        for (int s = 0; s < nsteps; ++s) {
            if (threads > 1)
                sim.advance_parallel(threads);
            else
                sim.advance();
        }
        store<T, D>(sim.particles(), x, v, F, C, Jp);
        store_grid<T, D>(sim.grid(), grid_v, grid_m);
    });
}

template <typename T, int D>
int p2g(int material, int res, T dt, T E, T nu, int idm, int64_t n, const T* x, const T* v, const T* F,
        const T* C, const T* Jp, const T* mass, const T* volume, T* grid_mom, T* grid_m) {
    return guarded([&] {
        oracle::Simulation<T, D> sim(load<T, D>(n, x, v, F, C, Jp, mass, volume), material, res, dt, E, nu, T(0),
                                     idm);
        sim.p2g();
        store_grid<T, D>(sim.grid(), grid_mom, grid_m);
    });
}

template <typename T, int D>
int grid_op(int res, T dt, T g, T* grid_v, T* grid_m) {
    return guarded([&] {
        oracle::Simulation<T, D> sim({}, oracle::kJelly, res, dt, T(1e4), T(0.2), g, 0);
        auto& cells = sim.grid_mut();
        cells.assign(size_t(sim.num_nodes()), oracle::Cell<T, D>{});
        load_grid<T, D>(cells, grid_v, grid_m);
        sim.grid_op();
        store_grid<T, D>(sim.grid(), grid_v, grid_m);
    });
}

template <typename T, int D>
int g2p(int material, int res, T dt, int idm, int64_t n, const T* grid_v, T* x, T* v, T* F, T* C, T* Jp) {
    return guarded([&] {
        std::vector<T> ones(static_cast<size_t>(n), T(1));
        oracle::Simulation<T, D> sim(load<T, D>(n, x, v, F, C, Jp, ones.data(), ones.data()), material, res, dt,
                                     T(1e4), T(0.2), T(0), idm);
        auto& cells = sim.grid_mut();
        cells.assign(size_t(sim.num_nodes()), oracle::Cell<T, D>{});
        std::vector<T> zeros(cells.size(), T(0));
        load_grid<T, D>(cells, grid_v, zeros.data());
        sim.g2p();
        store<T, D>(sim.particles(), x, v, F, C, Jp);
    });
}

template <typename T, int D>
void svd_batch(int64_t n, const T* A, T* U, T* S, T* V) {
    for (int64_t i = 0; i < n; ++i) {
        T a[D][D], u[D][D], s[D], vv[D][D];
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) a[r][c] = A[(i * D + r) * D + c];
        oracle::nclr_svd<T, D>(a, u, s, vv);
        for (int r = 0; r < D; ++r) {
            S[i * D + r] = s[r];
            for (int c = 0; c < D; ++c) {
                U[(i * D + r) * D + c] = u[r][c];
                V[(i * D + r) * D + c] = vv[r][c];
            }
        }
    }
}

template <typename T, int D>
void polar_batch(int64_t n, const T* A, T* R) {
    for (int64_t i = 0; i < n; ++i) {
        T a[D][D], r[D][D];
        for (int p = 0; p < D; ++p)
            for (int c = 0; c < D; ++c) a[p][c] = A[(i * D + p) * D + c];
        oracle::nclr_polar<T, D>(a, r);
        for (int p = 0; p < D; ++p)
            for (int c = 0; c < D; ++c) R[(i * D + p) * D + c] = r[p][c];
    }
}

template <typename T, int D>
void stress_batch(int material, int res, T dt, T E, T nu, int64_t n, const T* F, const T* C, const T* Jp,
                  const T* mass, const T* volume, T* out) {
    std::vector<oracle::Particle<T, D>> ps(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        auto& p = ps[size_t(i)];
        for (int a = 0; a < D; ++a) {
            p.x[a] = T(0.5);
            p.v[a] = T(0);
            for (int b = 0; b < D; ++b) {
                p.F[a][b] = F[(i * D + a) * D + b];
                p.C[a][b] = C[(i * D + a) * D + b];
            }
        }
        p.Jp = Jp[i];
        p.mass = mass[i];
        p.volume = volume[i];
    }
    oracle::Simulation<T, D> sim(ps, material, res, dt, E, nu, T(0), 0);
    for (int64_t i = 0; i < n; ++i) {
        T s[D][D];
        sim.first_piola_kirchoff_stress(sim.particles()[size_t(i)], s);
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) out[(i * D + a) * D + b] = s[a][b];
    }
}

template <int D>
int adjoint(int material, int res, double dt, double E, double nu, double g, int idm, int64_t n, const double* x,
            const double* v, const double* F, const double* C, const double* Jp, const double* mass,
            const double* volume, const double* gx, const double* gv, const double* gF, const double* gC,
            const double* gJ, double* ix, double* iv, double* iF, double* iC, double* iJ, double* iE) {
    return guarded([&] {
        const auto ps = load<double, D>(n, x, v, F, C, Jp, mass, volume);
        oracle::StepGrads<double, D> out(static_cast<size_t>(n)), in;
        std::copy(gx, gx + n * D, out.x.begin());
        std::copy(gv, gv + n * D, out.v.begin());
        std::copy(gF, gF + n * D * D, out.F.begin());
        std::copy(gC, gC + n * D * D, out.C.begin());
        if (gJ) std::copy(gJ, gJ + n, out.Jp.begin());
        oracle::adjoint_step<double, D>(ps, material, res, dt, E, nu, g, idm, out, in);
        std::copy(in.x.begin(), in.x.end(), ix);
        std::copy(in.v.begin(), in.v.end(), iv);
        std::copy(in.F.begin(), in.F.end(), iF);
        std::copy(in.C.begin(), in.C.end(), iC);
        if (iJ) std::copy(in.Jp.begin(), in.Jp.end(), iJ);
        iE[0] = in.E;
        iE[1] = in.E_abs;
    });
}

template <int D>
void polar_vjp(int64_t n, const double* A, const double* Rbar, double* Abar) {
    for (int64_t i = 0; i < n; ++i) {
        double a[D][D], r[D][D], rb[D][D], fb[D][D] = {};
        for (int p = 0; p < D; ++p)
            for (int c = 0; c < D; ++c) {
                a[p][c] = A[(i * D + p) * D + c];
                rb[p][c] = Rbar[(i * D + p) * D + c];
            }
        oracle::nclr_polar<double, D>(a, r);
        oracle::polar_backward<double, D>(a, r, rb, fb);
        for (int p = 0; p < D; ++p)
            for (int c = 0; c < D; ++c) Abar[(i * D + p) * D + c] = fb[p][c];
    }
}

}  // namespace

#define DIM_DISPATCH(dim, call2, call3) \
    ((dim) == 2 ? (call2) : (dim) == 3 ? (call3) : (g_err = "dim must be 2 or 3", 3))

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

#define DEFINE_FOR(T, SUF)                                                                                          \
    int oracle_run_##SUF(int dim, int material, int res, T dt, T E, T nu, T g, int idm, int64_t n, int nsteps,      \
                         T* x, T* v, T* F, T* C, T* Jp, const T* mass, const T* volume, T* grid_v, T* grid_m) {    \
        return DIM_DISPATCH(dim,                                                                                    \
                            (run<T, 2>(material, res, dt, E, nu, g, idm, n, nsteps, x, v, F, C, Jp, mass, volume,  \
                                       grid_v, grid_m)),                                                            \
                            (run<T, 3>(material, res, dt, E, nu, g, idm, n, nsteps, x, v, F, C, Jp, mass, volume,  \
                                       grid_v, grid_m)));                                                           \
    }                                                                                                               \
    int oracle_p2g_##SUF(int dim, int material, int res, T dt, T E, T nu, int idm, int64_t n, const T* x,           \
                         const T* v, const T* F, const T* C, const T* Jp, const T* mass, const T* volume,           \
                         T* grid_mom, T* grid_m) {                                                                  \
        return DIM_DISPATCH(                                                                                        \
            dim, (p2g<T, 2>(material, res, dt, E, nu, idm, n, x, v, F, C, Jp, mass, volume, grid_mom, grid_m)),     \
            (p2g<T, 3>(material, res, dt, E, nu, idm, n, x, v, F, C, Jp, mass, volume, grid_mom, grid_m)));         \
    }                                                                                                               \
    int oracle_grid_op_##SUF(int dim, int res, T dt, T g, T* grid_v, T* grid_m) {                                   \
        return DIM_DISPATCH(dim, (grid_op<T, 2>(res, dt, g, grid_v, grid_m)),                                       \
                            (grid_op<T, 3>(res, dt, g, grid_v, grid_m)));                                           \
    }                                                                                                               \
    int oracle_g2p_##SUF(int dim, int material, int res, T dt, int idm, int64_t n, const T* grid_v, T* x, T* v,     \
                         T* F, T* C, T* Jp) {                                                                       \
        return DIM_DISPATCH(dim, (g2p<T, 2>(material, res, dt, idm, n, grid_v, x, v, F, C, Jp)),                    \
                            (g2p<T, 3>(material, res, dt, idm, n, grid_v, x, v, F, C, Jp)));                        \
    }                                                                                                               \
    int oracle_svd_##SUF(int dim, int64_t n, const T* A, T* U, T* S, T* V) {                                        \
        return DIM_DISPATCH(dim, (svd_batch<T, 2>(n, A, U, S, V), 0), (svd_batch<T, 3>(n, A, U, S, V), 0));         \
    }                                                                                                               \
    int oracle_polar_##SUF(int dim, int64_t n, const T* A, T* R) {                                                  \
        return DIM_DISPATCH(dim, (polar_batch<T, 2>(n, A, R), 0), (polar_batch<T, 3>(n, A, R), 0));                 \
    }                                                                                                               \
    int oracle_stress_##SUF(int dim, int material, int res, T dt, T E, T nu, int64_t n, const T* F, const T* C,     \
                            const T* Jp, const T* mass, const T* volume, T* out) {                                  \
        return DIM_DISPATCH(dim, (stress_batch<T, 2>(material, res, dt, E, nu, n, F, C, Jp, mass, volume, out), 0), \
                            (stress_batch<T, 3>(material, res, dt, E, nu, n, F, C, Jp, mass, volume, out), 0));     \
    }                                                                                                               \
    void oracle_linspaced_##SUF(int n, T low, T high, T* out) {                                                     \
        auto v = oracle::linspaced<T>(n, low, high);                                                                \
        for (int i = 0; i < n; ++i) out[i] = v[size_t(i)];                                                          \
    }

DEFINE_FOR(float, f32)
DEFINE_FOR(double, f64)

// FP64 vector-Jacobian product of one step of any material (nclr_adjoint.hpp): cotangents of
// x, v, F, C and Jp (gJ / iJ may be NULL for zero / not wanted); iE[0] = dL/dE, iE[1] = the sum
// of the particles' |contributions| to it.
int oracle_adjoint_f64(int dim, int material, int res, double dt, double E, double nu, double g, int idm, int64_t n,
                       const double* x, const double* v, const double* F, const double* C, const double* Jp,
                       const double* mass, const double* volume, const double* gx, const double* gv,
                       const double* gF, const double* gC, const double* gJ, double* ix, double* iv, double* iF,
                       double* iC, double* iJ, double* iE) {
    return DIM_DISPATCH(dim,
                        (adjoint<2>(material, res, dt, E, nu, g, idm, n, x, v, F, C, Jp, mass, volume, gx, gv, gF,
                                    gC, gJ, ix, iv, iF, iC, iJ, iE)),
                        (adjoint<3>(material, res, dt, E, nu, g, idm, n, x, v, F, C, Jp, mass, volume, gx, gv, gF,
                                    gC, gJ, ix, iv, iF, iC, iJ, iE)));
}

// FP64 cotangent of the polar rotation factor: Abar = d<Rbar, R(A)>/dA.
int oracle_polar_vjp_f64(int dim, int64_t n, const double* A, const double* Rbar, double* Abar) {
    return DIM_DISPATCH(dim, (polar_vjp<2>(n, A, Rbar, Abar), 0), (polar_vjp<3>(n, A, Rbar, Abar), 0));
}

// Times `steps` calls of advance() after `warmup` untimed ones on an FP32
// simulation built from the given particles (the reference's serial CPU
// path, restated).  Returns seconds, or a negative value on error.
// FP32 steps of the oracle on `threads` OpenMP threads (1: the serial step the reference runs).
int oracle_run_threads_f32(int dim, int material, int res, float dt, float E, float nu, float g, int idm, int64_t n,
                           int nsteps, int threads, float* x, float* v, float* F, float* C, float* Jp,
                           const float* mass, const float* volume) {
    return DIM_DISPATCH(dim,
                        (run<float, 2>(material, res, dt, E, nu, g, idm, n, nsteps, x, v, F, C, Jp, mass, volume,
                                       nullptr, nullptr, threads)),
                        (run<float, 3>(material, res, dt, E, nu, g, idm, n, nsteps, x, v, F, C, Jp, mass, volume,
                                       nullptr, nullptr, threads)));
}

// FP64 steps on `threads` OpenMP threads (bitwise the serial step).
int oracle_run_threads_f64(int dim, int material, int res, double dt, double E, double nu, double g, int idm,
                           int64_t n, int nsteps, int threads, double* x, double* v, double* F, double* C, double* Jp,
                           const double* mass, const double* volume) {
    return DIM_DISPATCH(dim,
                        (run<double, 2>(material, res, dt, E, nu, g, idm, n, nsteps, x, v, F, C, Jp, mass, volume,
                                        nullptr, nullptr, threads)),
                        (run<double, 3>(material, res, dt, E, nu, g, idm, n, nsteps, x, v, F, C, Jp, mass, volume,
                                        nullptr, nullptr, threads)));
}

double oracle_time_steps_f32(int dim, int material, int res, float dt, float E, float nu, float g, int idm,
                             int64_t n, int warmup, int steps, const float* x, const float* v, const float* F,
                             const float* C, const float* Jp, const float* mass, const float* volume, int threads) {
    double secs = -1.0;
    auto body = [&](auto dim_tag) {
        constexpr int D = decltype(dim_tag)::value;
        oracle::Simulation<float, D> sim(load<float, D>(n, x, v, F, C, Jp, mass, volume), material, res, dt, E, nu,
                                         g, idm);
        auto step = [&] {
            if (threads > 1)
                sim.advance_parallel(threads);
            else
                sim.advance();
        };
        for (int s = 0; s < warmup; ++s) step();
        const auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < steps; ++s) step();
        const auto t1 = std::chrono::steady_clock::now();
        secs = std::chrono::duration<double>(t1 - t0).count();
    };
    const int rc = guarded([&] {
        if (dim == 2)
            body(std::integral_constant<int, 2>{});
        else
            body(std::integral_constant<int, 3>{});
    });
    return rc == 0 ? secs : -1.0;
}

}  // extern "C"
