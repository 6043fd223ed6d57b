// The CPU oracle's C API (oracle/oracle_capi.cpp) compiled into one program and driven through
// every step, adjoint and decomposition entry point on small scenes, for a build with
// -fsanitize=address,undefined (tests/test_sanitizers.py): out-of-bounds grid and stencil
// accesses, signed overflow and other undefined behaviour abort the run.
//   oracle_sanitize        prints "ok" and exits 0 when every call returned 0
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../../oracle/oracle_capi.cpp"

namespace {

struct Scene {
    int dim;
    int64_t n;
    std::vector<double> x, v, F, C, Jp, mass, vol;
};

// A cube of side^dim particles in [lo, lo + 0.2]^dim moving at v0 (towards a wall for the
// boundary paths), F = I plus noise, C random, Jp = 1.
Scene make_scene(int dim, int side, double lo, double v0, unsigned seed) {
    std::mt19937 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    Scene s;
    s.dim = dim;
    s.n = 1;
    for (int a = 0; a < dim; ++a) s.n *= side;
    for (int64_t i = 0; i < s.n; ++i) {
        int64_t r = i;
        for (int a = 0; a < dim; ++a) {
            const int k = int(r % side);
            r /= side;
            s.x.push_back(lo + 0.2 * (k + 0.5) / side);
            s.v.push_back(v0 + 0.1 * nd(rng));
        }
        for (int a = 0; a < dim; ++a)
            for (int b = 0; b < dim; ++b) {
                s.F.push_back((a == b ? 1.0 : 0.0) + 0.05 * nd(rng));
                s.C.push_back(0.5 * nd(rng));
            }
        s.Jp.push_back(1.0);
        s.mass.push_back(1.0);
        s.vol.push_back(1.0);
    }
    return s;
}

template <typename T>
std::vector<T> as(const std::vector<double>& a) {
    return std::vector<T>(a.begin(), a.end());
}

int failures = 0;

void expect_ok(int rc, const char* what, int dim, int mat) {
    if (rc != 0) {
        std::printf("FAILED %s dim=%d material=%d: %s\n", what, dim, mat, oracle_last_error());
        ++failures;
    }
}

void steps(const Scene& s, int mat, int idm, int nsteps) {
    const int res = 32;
    {
        auto x = as<float>(s.x), v = as<float>(s.v), F = as<float>(s.F), C = as<float>(s.C), Jp = as<float>(s.Jp);
        const auto m = as<float>(s.mass), V = as<float>(s.vol);
        int64_t nodes = 1;
        for (int a = 0; a < s.dim; ++a) nodes *= res + 1;
        std::vector<float> gv(size_t(nodes) * s.dim), gm(static_cast<size_t>(nodes));
        expect_ok(oracle_run_f32(s.dim, mat, res, 1e-4f, 1e4f, 0.2f, -100.0f, idm, s.n, nsteps, x.data(), v.data(),
                                 F.data(), C.data(), Jp.data(), m.data(), V.data(), gv.data(), gm.data()),
                  "run_f32", s.dim, mat);
    }
    {
        auto x = as<float>(s.x), v = as<float>(s.v), F = as<float>(s.F), C = as<float>(s.C), Jp = as<float>(s.Jp);
        const auto m = as<float>(s.mass), V = as<float>(s.vol);
        expect_ok(oracle_run_threads_f32(s.dim, mat, res, 1e-4f, 1e4f, 0.2f, -100.0f, idm, s.n, nsteps, 4, x.data(),
                                         v.data(), F.data(), C.data(), Jp.data(), m.data(), V.data()),
                  "run_threads_f32", s.dim, mat);
    }
    auto x = s.x, v = s.v, F = s.F, C = s.C, Jp = s.Jp;
    expect_ok(oracle_run_f64(s.dim, mat, res, 1e-4, 1e4, 0.2, -100.0, idm, s.n, nsteps, x.data(), v.data(), F.data(),
                             C.data(), Jp.data(), s.mass.data(), s.vol.data(), nullptr, nullptr),
              "run_f64", s.dim, mat);
    // the adjoint of the next step from the state reached
    const size_t dd = size_t(s.dim) * s.dim;
    std::vector<double> gx(x.size(), 1.0), gvv(v.size(), -0.5), gF(F.size(), 0.25), gC(C.size(), 0.1),
        gJ(Jp.size(), 0.3);
    std::vector<double> ix(x.size()), iv(v.size()), iF(s.n * dd), iC(s.n * dd), iJ(Jp.size()), iE(2);
    expect_ok(oracle_adjoint_f64(s.dim, mat, res, 1e-4, 1e4, 0.2, -100.0, idm, s.n, x.data(), v.data(), F.data(),
                                 C.data(), Jp.data(), s.mass.data(), s.vol.data(), gx.data(), gvv.data(), gF.data(),
                                 gC.data(), gJ.data(), ix.data(), iv.data(), iF.data(), iC.data(), iJ.data(),
                                 iE.data()),
              "adjoint_f64", s.dim, mat);
}

void decompositions(int dim) {
    // random, singular, rank-1, zero and reflected matrices
    std::mt19937 rng(7 + dim);
    std::normal_distribution<double> nd(0.0, 1.0);
    const int64_t n = 64;
    const size_t dd = size_t(dim) * dim;
    std::vector<double> A(n * dd);
    for (int64_t i = 0; i < n; ++i)
        for (size_t k = 0; k < dd; ++k) A[i * dd + k] = nd(rng);
    for (size_t k = 0; k < dd; ++k) {
        A[0 * dd + k] = 0.0;                                       // zero
        A[1 * dd + k] = (k / dim == 0) ? 1.0 : 0.0;                // rank 1
        A[2 * dd + k] = (k % (dim + 1) == 0 && k < dd - 1) ? 1.0 : 0.0;  // diag(1, .., 1, 0)
        A[3 * dd + k] = (k % (dim + 1) == 0) ? (k == 0 ? -2.0 : 2.0) : 0.0;  // scaled reflection
    }
    std::vector<double> U(n * dd), S(n * dim), V(n * dd), R(n * dd), Rbar(n * dd, 0.5), Abar(n * dd);
    expect_ok(oracle_svd_f64(dim, n, A.data(), U.data(), S.data(), V.data()), "svd_f64", dim, -1);
    expect_ok(oracle_polar_f64(dim, n, A.data(), R.data()), "polar_f64", dim, -1);
    expect_ok(oracle_polar_vjp_f64(dim, n, A.data(), Rbar.data(), Abar.data()), "polar_vjp_f64", dim, -1);
    auto Af = as<float>(A);
    std::vector<float> Uf(n * dd), Sf(n * dim), Vf(n * dd);
    expect_ok(oracle_svd_f32(dim, n, Af.data(), Uf.data(), Sf.data(), Vf.data()), "svd_f32", dim, -1);
}

}  // namespace

int main() {
    for (int dim = 2; dim <= 3; ++dim) {
        decompositions(dim);
        const int side = dim == 2 ? 12 : 5;
        for (int mat = 0; mat < 3; ++mat)
            for (int idm = 0; idm < 2; ++idm) {
                steps(make_scene(dim, side, 0.4, 0.0, 11u + unsigned(mat)), mat, idm, 4);
                // flying into the low wall: the sticky-boundary and clamp paths
                steps(make_scene(dim, side, 0.04, -60.0, 23u + unsigned(mat)), mat, idm, 6);
            }
    }
    // a particle outside the grid: the step reports out_of_range through the error path
    for (int dim = 2; dim <= 3; ++dim) {
        Scene s = make_scene(dim, 3, 0.4, 0.0, 5u);
        s.x[0] = 1.5;
        auto x = s.x, v = s.v, F = s.F, C = s.C, Jp = s.Jp;
        const int rc = oracle_run_f64(dim, 0, 16, 1e-4, 1e4, 0.2, -100.0, 0, s.n, 1, x.data(), v.data(), F.data(),
                                      C.data(), Jp.data(), s.mass.data(), s.vol.data(), nullptr, nullptr);
        if (rc == 0 || oracle_last_error()[0] == '\0') {
            std::printf("FAILED out-of-range particle not reported (dim=%d)\n", dim);
            ++failures;
        }
    }
    std::vector<double> lin(17);
    oracle_linspaced_f64(17, -0.3, 0.7, lin.data());
    std::printf(failures ? "failures %d\n" : "ok\n", failures);
    return failures ? 1 : 0;
}
