// Test infrastructure only: the CPU oracle for the MLS-MPM step.
//
// This header restates, in plain C++17 with no Eigen, the algorithm of the
// reference engine's step `MPMSimulation<dim>::advance()` (src/nclr.h:80-84)
// and the small-matrix helpers it calls (src/nclr_math.h:13-19, 40-129).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
// and only as the checker; the product library never links it.
//
// It is templated on the scalar type T: T=float follows the reference's FP32
// arithmetic (nclr_math.h:11 `using real = float`) including the places where
// the reference silently widens to double (snow hardening exp, liquid J,
// allowed_velocity), and T=double is the FP64 oracle used for drift reports
// and finite-difference gradient checks.
//
// Operation order inside each formula follows the reference's expression
// structure left to right (e.g. `2*mu*(F-r)*F^T` is ((2mu)(F-r))F^T).  The
// reference is built with -Ofast (CMakeLists.txt:9-10), so its own last-bit
// rounding is not reproducible and is not claimed here.
//
// Pinning (what checks each function; see tests/test_oracle_*.py):
//   jacobi_svd / nclr_svd  : numpy FP64 linalg.svd, U S V^T = A, U,V in SO(3),
//                            sorted |sigma|, sign on sigma_min (3D)
//   polar                  : R in SO(d), R^T A symmetric positive, numpy polar
//   weights                : partition of unity, first moment zero
//   p2g                    : mass / momentum conservation, FP64 brute force
//   grid_op                : closed form per node, sticky box edge cases
//   g2p                    : affine reproduction (uniform + linear fields),
//                            snow clamp vs numpy SVD, Jp and liquid J closed forms
//   cube                   : lattice spacing / endpoints
// Parity unpinned (no reference build is possible: Eigen3 is absent):
//   the exact U/V choice of Eigen's JacobiSVD for rank-deficient inputs and
//   the undefined behaviour of nclr_svd's col(2) flip in 2D (treated here as
//   a no-op, sigma >= 0).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <exception>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

enum Material : int { kSnow = 0, kJelly = 1, kLiquid = 2 };

// identity_mode 0 = reference mode: diag<dim>(1) only writes (0,0) and (1,1)
// (nclr_math.h:13-19), so in 3D the "identity" is diag(1,1,0).
// identity_mode 1 = physical identity.
template <typename T, int D>
inline void diag_ref(T out[D][D], T value, int identity_mode) {
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) out[i][j] = T(0);
    out[0][0] = value;
    out[1][1] = value;
    if (identity_mode == 1)
        for (int i = 2; i < D; ++i) out[i][i] = value;
}

template <typename T>
inline T det(const T a[2][2]) {
    return a[0][0] * a[1][1] - a[0][1] * a[1][0];
}

// Cofactor expansion along the first row.
template <typename T>
inline T det(const T a[3][3]) {
    return a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
           a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
           a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
}

template <typename T, int D>
inline void matmul(const T a[D][D], const T b[D][D], T out[D][D]) {
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            T s = a[i][0] * b[0][j];
            for (int k = 1; k < D; ++k) s = s + a[i][k] * b[k][j];
            out[i][j] = s;
        }
}

// out = a * b^T
template <typename T, int D>
inline void matmul_bt(const T a[D][D], const T b[D][D], T out[D][D]) {
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            T s = a[i][0] * b[j][0];
            for (int k = 1; k < D; ++k) s = s + a[i][k] * b[j][k];
            out[i][j] = s;
        }
}

// ---------------------------------------------------------------------------
// SVD: two-sided cyclic Jacobi with a symmetrising pre-rotation per 2x2 block.
// This restates the published algorithm of Eigen 3.4's JacobiSVD for square
// fixed-size real matrices, which nclr_svd calls (nclr_math.h:57); Eigen is
// an unpinned system dependency of the reference (CMakeLists.txt:12).
// Steps: scale by max|a_ij|; sweep pairs (p,q) = (1,0),(2,0),(2,1); for each
// pair whose off-diagonals exceed max(min_normal, 2*eps*max|diag|) compute a
// left rotation that symmetrises the 2x2 block and a Jacobi rotation that
// diagonalises it (Golub & Van Loan, sym. Schur 2x2); accumulate into U and V;
// make the diagonal non-negative (flipping U columns); unscale; selection-sort
// descending.
// ---------------------------------------------------------------------------
template <typename T>
struct Rot {
    T c, s;
};

// Rows p,q of W <- [[c, s], [-s, c]] * rows p,q
template <typename T, int D>
inline void rot_rows(T W[D][D], int p, int q, Rot<T> r) {
    for (int k = 0; k < D; ++k) {
        const T x = W[p][k], y = W[q][k];
        W[p][k] = r.c * x + r.s * y;
        W[q][k] = -r.s * x + r.c * y;
    }
}

// Columns p,q of M <- cols p,q * [[c, s], [-s, c]]
template <typename T, int D>
inline void rot_cols(T M[D][D], int p, int q, Rot<T> r) {
    for (int k = 0; k < D; ++k) {
        const T x = M[k][p], y = M[k][q];
        M[k][p] = r.c * x - r.s * y;
        M[k][q] = r.s * x + r.c * y;
    }
}

// Columns p,q of M <- cols p,q * [[c, s], [-s, c]]^T
template <typename T, int D>
inline void rot_cols_t(T M[D][D], int p, int q, Rot<T> r) {
    for (int k = 0; k < D; ++k) {
        const T x = M[k][p], y = M[k][q];
        M[k][p] = r.c * x + r.s * y;
        M[k][q] = -r.s * x + r.c * y;
    }
}

// Jacobi rotation that diagonalises the symmetric 2x2 [[x, y], [y, z]].
template <typename T>
inline Rot<T> make_jacobi(T x, T y, T z) {
    const T deno = T(2) * std::fabs(y);
    if (deno < std::numeric_limits<T>::min()) return {T(1), T(0)};
    const T tau = (x - z) / deno;
    const T w = std::sqrt(tau * tau + T(1));
    const T t = tau > T(0) ? T(1) / (tau + w) : T(1) / (tau - w);
    const T sign_t = t > T(0) ? T(1) : T(-1);
    const T n = T(1) / std::sqrt(t * t + T(1));
    const T s = -sign_t * (y / std::fabs(y)) * std::fabs(t) * n;
    return {n, s};
}

template <typename T, int D>
inline void svd_2x2_block(const T W[D][D], int p, int q, Rot<T>* left, Rot<T>* right) {
    T m00 = W[p][p], m01 = W[p][q], m10 = W[q][p], m11 = W[q][q];
    Rot<T> rot1;
    const T t = m00 + m11;
    const T d = m10 - m01;
    if (std::fabs(d) < std::numeric_limits<T>::min()) {
        rot1 = {T(1), T(0)};
    } else {
        const T u = t / d;
        const T tmp = std::sqrt(T(1) + u * u);
        rot1 = {u / tmp, T(1) / tmp};
    }
    // symmetrise: rows (0,1) of m <- rot1 * m
    const T a00 = rot1.c * m00 + rot1.s * m10;
    const T a01 = rot1.c * m01 + rot1.s * m11;
    const T a11 = -rot1.s * m01 + rot1.c * m11;
    const Rot<T> jr = make_jacobi(a00, a01, a11);
    *right = jr;
    // left = rot1 * right^T  (composition of plane rotations)
    const T rc = jr.c, rs = -jr.s;
    *left = {rot1.c * rc - rot1.s * rs, rot1.c * rs + rot1.s * rc};
}

constexpr int kMaxJacobiSweeps = 32;

template <typename T, int D>
inline void jacobi_svd(const T A[D][D], T U[D][D], T sig[D], T V[D][D]) {
    T scale = T(0);
    bool finite = true;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            if (!std::isfinite(A[i][j])) finite = false;
            scale = std::max(scale, std::fabs(A[i][j]));
        }
    if (!finite) {
        const T nan = std::numeric_limits<T>::quiet_NaN();
        for (int i = 0; i < D; ++i) {
            sig[i] = nan;
            for (int j = 0; j < D; ++j) U[i][j] = V[i][j] = nan;
        }
        return;
    }
    if (scale == T(0)) scale = T(1);
    T W[D][D];
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) {
            W[i][j] = A[i][j] / scale;
            U[i][j] = V[i][j] = (i == j) ? T(1) : T(0);
        }
    T max_diag = T(0);
    for (int i = 0; i < D; ++i) max_diag = std::max(max_diag, std::fabs(W[i][i]));
    const T consider_zero = std::numeric_limits<T>::min();
    const T precision = T(2) * std::numeric_limits<T>::epsilon();
    bool finished = false;
    for (int sweep = 0; !finished && sweep < kMaxJacobiSweeps; ++sweep) {
        finished = true;
        for (int p = 1; p < D; ++p) {
            for (int q = 0; q < p; ++q) {
                const T threshold = std::max(consider_zero, precision * max_diag);
                if (std::fabs(W[p][q]) > threshold || std::fabs(W[q][p]) > threshold) {
                    finished = false;
                    Rot<T> jl, jr;
                    svd_2x2_block<T, D>(W, p, q, &jl, &jr);
                    rot_rows<T, D>(W, p, q, jl);
                    rot_cols_t<T, D>(U, p, q, jl);
                    rot_cols<T, D>(W, p, q, jr);
                    rot_cols<T, D>(V, p, q, jr);
                    max_diag = std::max(max_diag, std::max(std::fabs(W[p][p]), std::fabs(W[q][q])));
                }
            }
        }
    }
    for (int i = 0; i < D; ++i) {
        const T a = W[i][i];
        sig[i] = std::fabs(a);
        if (a < T(0))
            for (int k = 0; k < D; ++k) U[k][i] = -U[k][i];
    }
    for (int i = 0; i < D; ++i) sig[i] = sig[i] * scale;
    for (int i = 0; i < D; ++i) {
        int pos = i;
        for (int k = i + 1; k < D; ++k)
            if (sig[k] > sig[pos]) pos = k;
        if (sig[pos] == T(0)) break;
        if (pos != i) {
            std::swap(sig[i], sig[pos]);
            for (int k = 0; k < D; ++k) {
                std::swap(U[k][i], U[k][pos]);
                std::swap(V[k][i], V[k][pos]);
            }
        }
    }
}

// nclr_svd (nclr_math.h:50-74): Jacobi SVD, then in 3D flip column 2 of U
// (resp. V) and the sign of sigma_2 when det U < 0 (resp. det V < 0), giving
// U, V in SO(3) with the sign of det A on the smallest singular value.  In 2D
// the reference's col(2) access is out of bounds (undefined behaviour); this
// restatement treats it as a no-op, so 2D singular values stay >= 0.
template <typename T, int D>
inline void nclr_svd(const T A[D][D], T U[D][D], T sig[D], T V[D][D]) {
    jacobi_svd<T, D>(A, U, sig, V);
    if constexpr (D == 3) {
        if (det<T>(U) < T(0)) {
            for (int k = 0; k < 3; ++k) U[k][2] = -U[k][2];
            sig[2] = -sig[2];
        }
        if (det<T>(V) < T(0)) {
            for (int k = 0; k < 3; ++k) V[k][2] = -V[k][2];
            sig[2] = -sig[2];
        }
    }
}

// nclr_polar (nclr_math.h:76-98): rotation factor R of A = R S.
// 2D: closed form from (a00 + a11, a10 - a01).  3D: R = U V^T from nclr_svd.
template <typename T, int D>
inline void nclr_polar(const T A[D][D], T R[D][D]) {
    if constexpr (D == 2) {
        const T x = A[0][0] + A[1][1];
        const T y = A[1][0] - A[0][1];
        const T scale = T(1) / std::sqrt(x * x + y * y);
        const T c = x * scale, s = y * scale;
        R[0][0] = c;
        R[0][1] = -s;
        R[1][0] = s;
        R[1][1] = c;
    } else {
        T U[3][3], sig[3], V[3][3];
        nclr_svd<T, 3>(A, U, sig, V);
        matmul_bt<T, 3>(U, V, R);
    }
}

// Quadratic B-spline weights (nclr.h:124-127): for fx in [0.5, 1.5) per axis.
template <typename T>
inline void bspline_weights(T fx, T w[3]) {
    const T a = T(1.5) - fx;
    const T b = fx - T(1.0);
    const T c = fx - T(0.5);
    w[0] = T(0.5) * (a * a);
    w[1] = T(0.75) - b * b;
    w[2] = T(0.5) * (c * c);
}

// ---------------------------------------------------------------------------
// Simulation state (nclr.h:20-55)
// ---------------------------------------------------------------------------
template <typename T, int D>
struct Particle {
    T x[D];
    T v[D];
    T F[D][D];
    T C[D][D];
    T Jp;
    T mass;
    T volume;
    int c;
};

template <typename T, int D>
struct Cell {
    T v[D];
    T m;
};

struct OutOfRange : std::out_of_range {
    using std::out_of_range::out_of_range;
};

template <typename T, int D>
class Simulation {
  public:
    static constexpr int kBoundary = 3;
    // The reference declares these as `real` (float); T=double widens them.
    static constexpr T kSnowHardening = T(10.0f);
    static constexpr T kJellyHardening = T(0.3f);
    static constexpr T kLiquidHardening = T(1.0f);

    // nclr.h:74-78.  dx = 1.0/res is computed in double then narrowed.
    Simulation(std::vector<Particle<T, D>> particles, int material, int res, T dt, T E, T nu, T gravity,
               int identity_mode = 0)
        : particles_(std::move(particles)),
          material_(material),
          res_(res),
          dt_(dt),
          dx_(T(1.0 / res)),
          inv_dx_(T(1) / dx_),
          gravity_(gravity),
          identity_mode_(identity_mode),
          mu_0(E / (T(2) * (T(1) + nu))),
          lambda_0(E * nu / ((T(1) + nu) * (T(1) - T(2) * nu))) {}

    void advance() {
        p2g();
        grid_op();
        g2p();
    }

    std::vector<Particle<T, D>>& particles() { return particles_; }
    const std::vector<Cell<T, D>>& grid() const { return cells_; }
    std::vector<Cell<T, D>>& grid_mut() { return cells_; }
    int res() const { return res_; }
    T dt() const { return dt_; }
    T dx() const { return dx_; }
    T inv_dx() const { return inv_dx_; }
    T gravity() const { return gravity_; }
    int material() const { return material_; }
    int identity_mode() const { return identity_mode_; }
    T allowed_velocity() const { return T(double(dx_) * 0.9 / double(dt_)); }

    // Flat index of node base+off, checked like cell_in (for callers outside the class).
    int64_t node_index(const int base[D], const int off[D]) const {
        int64_t idx = 0;
        for (int a = 0; a < D; ++a) idx = idx * (res_ + 1) + int64_t(base[a] + off[a]);
        if (idx < 0 || idx >= num_nodes()) throw OutOfRange("grid index out of range");
        return idx;
    }

    template <typename Fn>
    void each_offset(Fn&& fn) const {
        for_each_offset(std::forward<Fn>(fn));
    }

    int64_t num_nodes() const {
        int64_t n = 1;
        for (int a = 0; a < D; ++a) n *= (res_ + 1);
        return n;
    }

    // nclr.h:104-165.  Grid reallocated (zeroed) each call.
    void p2g() {
        zero_grid(cells_);
        for (size_t pp = 0; pp < particles_.size(); ++pp) deposit(particles_[pp], cells_);
    }

    // One particle's 3^D contributions (nclr.h:110-158) added into `grid`.
    void deposit(const Particle<T, D>& p, std::vector<Cell<T, D>>& grid) const {
        Prologue q;
        prologue(p, q);
        scatter(q, grid, std::numeric_limits<int>::min(), std::numeric_limits<int>::max());
    }

    // What a particle's deposit needs besides the grid: its stencil and fused stress (nclr.h:112-137).
    struct Prologue {
        int base[D];
        T fx[D], w[D][3];
        T affine[D][D];
        T mv[D];
        T mass;
    };

    void prologue(const Particle<T, D>& p, Prologue& q) const {
        base_and_weights(p.x, q.base, q.fx, q.w);
        first_piola_kirchoff_stress(p, q.affine);
        for (int a = 0; a < D; ++a) q.mv[a] = p.v[a] * p.mass;
        q.mass = p.mass;
    }

    // The contributions of one prologue to the nodes whose first index lies in [row_lo, row_hi)
    // (nclr.h:138-165), in the reference's offset order.
    void scatter(const Prologue& q, std::vector<Cell<T, D>>& grid, int row_lo, int row_hi) const {
        for_each_offset([&](const int off[D]) {
            const int row = q.base[0] + off[0];
            if (row < row_lo || row >= row_hi) return;
            T dpos[D];
            for (int a = 0; a < D; ++a) dpos[a] = (T(off[a]) - q.fx[a]) * dx_;
            T weight = q.w[0][off[0]];
            for (int a = 1; a < D; ++a) weight = weight * q.w[a][off[a]];
            Cell<T, D>& cell = cell_in(grid, q.base, off);
            for (int r = 0; r < D; ++r) {
                T ad = q.affine[r][0] * dpos[0];
                for (int c = 1; c < D; ++c) ad = ad + q.affine[r][c] * dpos[c];
                cell.v[r] = cell.v[r] + weight * (q.mv[r] + ad);
            }
            cell.m = cell.m + weight * q.mass;
        });
    }

    // nclr.h:263-310
    void grid_op() {
        const int64_t total = num_nodes();
        for (int64_t idx = 0; idx < total; ++idx) update_node(idx);
    }

    void update_node(int64_t idx) {
        const T allowed = T(double(dx_) * 0.9 / double(dt_));
        int n[D];
        int64_t rem = idx;
        for (int a = D - 1; a >= 0; --a) {
            n[a] = int(rem % (res_ + 1));
            rem /= (res_ + 1);
        }
        Cell<T, D>& g = cells_[idx];
        if (g.m > T(0)) {
            for (int a = 0; a < D; ++a) g.v[a] = g.v[a] / g.m;
            g.v[1] = g.v[1] + dt_ * gravity_;
            for (int a = 0; a < D; ++a) g.v[a] = std::clamp(g.v[a], -allowed, allowed);
        }
        for (int a = 0; a < D; ++a) {
            if ((n[a] < kBoundary && g.v[a] < T(0)) || (n[a] >= (res_ + 1) - kBoundary && g.v[a] > T(0))) {
                for (int b = 0; b < D; ++b) g.v[b] = T(0);
                g.m = T(0);
            }
        }
    }

    // The same step on `threads` OpenMP threads, bitwise equal to advance(): the per-particle
    // prologues (stencil and stress, the costly part) in parallel; then each thread owns a band of
    // first-axis node rows and adds, in particle order, the contributions that land in its band,
    // so every node receives its contributions in the serial order; then the grid update over
    // nodes and G2P over particles, both per element.  The first and last bands are unbounded, so
    // an out-of-range node still throws (std::vector::at in the reference).
    void advance_parallel(int threads) {
        threads = std::max(threads, 1);
        const int64_t nn = num_nodes();
        const int64_t np = int64_t(particles_.size());
        std::vector<Prologue>& pro = prologues_;  // kept between steps (no reallocation)
        pro.resize(size_t(np));
        std::exception_ptr err = nullptr;
        auto guard = [&](auto&& fn) {
            try {
                fn();
            } catch (...) {
#pragma omp critical
                if (!err) err = std::current_exception();
            }
        };
        zero_grid(cells_);
#pragma omp parallel num_threads(threads)
        {
#pragma omp for schedule(static)
            for (int64_t pp = 0; pp < np; ++pp) guard([&] { prologue(particles_[size_t(pp)], pro[size_t(pp)]); });
            int t = 0, nt = 1;
#ifdef _OPENMP
            t = omp_get_thread_num();
            nt = omp_get_num_threads();
#endif
            const int rows = res_ + 1;
            const int lo = t == 0 ? std::numeric_limits<int>::min() : int(int64_t(rows) * t / nt);
            const int hi = t == nt - 1 ? std::numeric_limits<int>::max() : int(int64_t(rows) * (t + 1) / nt);
            guard([&] {
                for (int64_t pp = 0; pp < np; ++pp) {
                    const Prologue& q = pro[size_t(pp)];
                    if (q.base[0] + 2 >= lo && q.base[0] < hi) scatter(q, cells_, lo, hi);
                }
            });
#pragma omp barrier
#pragma omp for schedule(static)
            for (int64_t i = 0; i < nn; ++i) update_node(i);
#pragma omp for schedule(static)
            for (int64_t pp = 0; pp < np; ++pp) guard([&] { update_particle(particles_[size_t(pp)]); });
        }
        if (err) std::rethrow_exception(err);
    }

    // nclr.h:167-261
    void g2p() {
        for (size_t pp = 0; pp < particles_.size(); ++pp) update_particle(particles_[pp]);
    }

    void update_particle(Particle<T, D>& p) const {
        {
            int base[D];
            T fx[D], w[D][3];
            base_and_weights(p.x, base, fx, w);
            for (int a = 0; a < D; ++a) {
                p.v[a] = T(0);
                for (int b = 0; b < D; ++b) p.C[a][b] = T(0);
            }
            const T four_inv_dx = T(4) * inv_dx_;
            for_each_offset([&](const int off[D]) {
                T dpos[D];
                for (int a = 0; a < D; ++a) dpos[a] = T(off[a]) - fx[a];
                const Cell<T, D>& cell = cell_in(cells_, base, off);
                T weight = w[0][off[0]];
                for (int a = 1; a < D; ++a) weight = weight * w[a][off[a]];
                for (int a = 0; a < D; ++a) p.v[a] = p.v[a] + weight * cell.v[a];
                for (int a = 0; a < D; ++a) {
                    const T wv = four_inv_dx * (weight * cell.v[a]);
                    for (int b = 0; b < D; ++b) p.C[a][b] = p.C[a][b] + wv * dpos[b];
                }
            });
            for (int a = 0; a < D; ++a) p.x[a] = p.x[a] + dt_ * p.v[a];
            T Fh[D][D], M[D][D];
            diag_ref<T, D>(M, T(1), identity_mode_);
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b) M[a][b] = M[a][b] + dt_ * p.C[a][b];
            matmul<T, D>(M, p.F, Fh);
            if (material_ == kJelly) {
                copy(Fh, p.F);
            } else {
                T U[D][D], sig[D], V[D][D];
                nclr_svd<T, D>(Fh, U, sig, V);
                if (material_ == kSnow) {
                    const T lo = T(float(1.0 - 2.5e-2));
                    const T hi = T(float(1.0 + 4.5e-3));
                    for (int d = 0; d < D; ++d) sig[d] = std::clamp(sig[d], lo, hi);
                    const T old_J = det<T>(Fh);
                    T US[D][D], Fn[D][D];
                    for (int a = 0; a < D; ++a)
                        for (int b = 0; b < D; ++b) US[a][b] = U[a][b] * sig[b];
                    matmul_bt<T, D>(US, V, Fn);
                    p.Jp = std::clamp(p.Jp * old_J / det<T>(Fn), T(0.6f), T(20.0f));
                    copy(Fn, p.F);
                }
                if (material_ == kLiquid) {
                    double J = 1.0;
                    for (int d = 0; d < D; ++d) J *= double(sig[d]);
                    diag_ref<T, D>(p.F, T(1), identity_mode_);
                    p.F[0][0] = T(J);
                }
            }
        }
    }

    // nclr.h:313-337: returns the fused APIC + MLS stress matrix
    //   -(dt V0)(4 inv_dx^2 PF) + m C,   PF = 2mu(F-R)F^T + lambda(J-1)J * ones
    void first_piola_kirchoff_stress(const Particle<T, D>& p, T out[D][D]) const {
        T mu, lambda;
        hardening(p, &mu, &lambda);
        const T J = det<T>(p.F);
        T R[D][D];
        nclr_polar<T, D>(p.F, R);
        const T Dinv = T(4) * inv_dx_ * inv_dx_;
        T FmR[D][D], PF[D][D];
        const T two_mu = T(2) * mu;
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) FmR[a][b] = two_mu * (p.F[a][b] - R[a][b]);
        matmul_bt<T, D>(FmR, p.F, PF);
        const T lam_term = lambda * (J - T(1)) * J;
        const T neg_dtv = -(dt_ * p.volume);
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b) {
                const T pf = PF[a][b] + lam_term;
                out[a][b] = neg_dtv * (Dinv * pf) + p.mass * p.C[a][b];
            }
    }

    // nclr.h:351-372.  Snow: e = exp(10 (1 - Jp)) evaluated in double.
    void hardening(const Particle<T, D>& p, T* mu, T* lambda) const {
        T e;
        switch (material_) {
            case kSnow:
                e = T(std::exp(double(kSnowHardening) * (1.0 - double(p.Jp))));
                break;
            case kJelly:
                e = kJellyHardening;
                break;
            case kLiquid:
                e = kLiquidHardening;
                break;
            default:
                *mu = T(0);
                *lambda = T(0);
                return;
        }
        *mu = mu_0 * e;
        *lambda = lambda_0 * e;
    }

    // nclr.h:115,121 (truncation toward zero, Q4) and 124-127.
    void base_and_weights(const T x[D], int base[D], T fx[D], T w[D][3]) const {
        for (int a = 0; a < D; ++a) {
            const T s = x[a] * inv_dx_;
            const T b = s - T(0.5);
            if (!std::isfinite(b) || b >= T(2147483647.0) || b <= T(-2147483648.0))
                throw OutOfRange("particle position is not a finite in-range number");
            base[a] = int(b);
            fx[a] = s - T(base[a]);
            bspline_weights<T>(fx[a], w[a]);
        }
    }

  private:
    std::vector<Prologue> prologues_;  // advance_parallel's per-particle prologues
    template <typename Fn>
    void for_each_offset(Fn&& fn) const {
        int off[D];
        if constexpr (D == 2) {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    off[0] = i;
                    off[1] = j;
                    fn(off);
                }
        } else {
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        off[0] = i;
                        off[1] = j;
                        off[2] = k;
                        fn(off);
                    }
        }
    }

    // Flattened x-major index, checked only on the flat range like
    // std::vector::at (nclr.h:141-142,152,163-164): per-axis overflow on the
    // faster axes aliases into a neighbouring row, as in the reference.
    template <typename Grid>
    auto cell_in(Grid& grid, const int base[D], const int off[D]) const -> decltype(grid[0]) {
        int64_t idx = 0;
        for (int a = 0; a < D; ++a) idx = idx * (res_ + 1) + int64_t(base[a] + off[a]);
        if (idx < 0 || idx >= int64_t(grid.size())) throw OutOfRange("grid index out of range");
        return grid[size_t(idx)];
    }

    void zero_grid(std::vector<Cell<T, D>>& grid) const {
        grid.assign(size_t(num_nodes()), Cell<T, D>{});
        for (auto& c : grid) {
            for (int a = 0; a < D; ++a) c.v[a] = T(0);
            c.m = T(0);
        }
    }

    static void copy(const T a[D][D], T b[D][D]) {
        for (int i = 0; i < D; ++i)
            for (int j = 0; j < D; ++j) b[i][j] = a[i][j];
    }

    std::vector<Particle<T, D>> particles_;
    std::vector<Cell<T, D>> cells_;
    int material_;
    int res_;
    T dt_;
    T dx_;
    T inv_dx_;
    T gravity_;
    int identity_mode_;

  public:
    const T mu_0;
    const T lambda_0;
};

// cube<dim>(res, min, max) (nclr_math.h:100-129): res^dim lattice, last axis
// fastest, coordinates from a LinSpaced(res, min, max) sequence:
//   value(i) = low + i*step, step = (high-low)/(res-1), last value = high
// (when |high| < |low| the sequence is built from the high end instead).
template <typename T>
inline std::vector<T> linspaced(int n, T low, T high) {
    std::vector<T> out(size_t(std::max(n, 0)));
    if (n <= 0) return out;
    if (n == 1) {
        out[0] = high;
        return out;
    }
    const int size1 = n - 1;
    const T step = (high - low) / T(size1);
    const bool flip = std::fabs(high) < std::fabs(low);
    for (int i = 0; i < n; ++i) {
        if (flip)
            out[i] = (i == 0) ? low : T(high - T(size1 - i) * step);
        else
            out[i] = (i == size1) ? high : T(low + T(i) * step);
    }
    return out;
}

}  // namespace oracle
