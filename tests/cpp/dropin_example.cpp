// Headless caller written against the drop-in API exactly the way the reference's
// src/example.cpp:29-53 and src/solver.cpp:133-188 use it (no GUI, no flags library).
//
//   dropin_example <steps> <out.txt>
// runs the two-cube 2D snow scene for <steps> steps and writes "n dim" then one line per
// particle (x, v, F, C, Jp) in the original particle order; exits non-zero on an API mismatch.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <memory>

#include "nclr.h"

constexpr int kDimension = 2;
constexpr int kResolution = 50;
constexpr int kColor = 0xED553B;

#define EXPECT(cond)                                                             \
    do {                                                                         \
        if (!(cond)) {                                                           \
            std::cerr << "dropin_example: check failed: " #cond << std::endl;    \
            return 3;                                                            \
        }                                                                        \
    } while (0)

int main(int argc, char** argv) {
    using namespace nclr;
    const int steps = argc > 1 ? std::atoi(argv[1]) : 10;
    const char* out_path = argc > 2 ? argv[2] : "dropin_out.txt";

    // --- src/example.cpp:29-46 ----------------------------------------------------------
    std::vector<Particle<kDimension>> particles;
    auto cube_particles = cube<kDimension>(kResolution, 0.4, 0.6);
    for (const auto& pos : cube_particles) { particles.emplace_back(Particle<kDimension>(pos, kColor)); }
    cube_particles = cube<kDimension>(kResolution, 0.4, 0.6);
    for (auto& pos : cube_particles) {
        pos(1) -= 0.35;
        particles.emplace_back(Particle<2>(pos, kColor));
    }
    auto sim = std::make_unique<MPMSimulation<kDimension>>(particles, MaterialModel::kSnow);
    EXPECT(sim->grid().empty());
    for (int frame = 0; frame < steps; ++frame) sim->advance();
    const auto& ps = sim->particles();
    EXPECT(ps.size() == 5000);
    EXPECT(ps[0].c == kColor);
    std::ofstream out(out_path);
    out.precision(9);
    out << ps.size() << " " << kDimension << "\n";
    for (const auto& p : ps) {
        out << p.x(0) << " " << p.x(1) << " " << p.v(0) << " " << p.v(1) << " " << p.F(0, 0) << " " << p.F(0, 1) << " "
            << p.F(1, 0) << " " << p.F(1, 1) << " " << p.C(0, 0) << " " << p.C(0, 1) << " " << p.C(1, 0) << " "
            << p.C(1, 1) << " " << p.Jp << "\n";
    }
    out.close();

    // --- src/solver.cpp:184-188, 51-53, 100-104 ------------------------------------------
    constexpr int kGridResolution = 64;
    constexpr nclr::real kDt = 1e-4;
    std::vector<nclr::Particle<2>> small;
    for (const auto& pos : nclr::cube<2>(25, 0.4, 0.6)) small.emplace_back(nclr::Particle<2>(pos, kColor));
    auto solver = std::make_unique<nclr::MPMSimulation<2>>(small, nclr::MaterialModel::kJelly, kGridResolution, kDt,
                                                           1000.0, 0.3, -100.0);
    solver->advance();
    const auto cells = solver->grid();
    EXPECT(cells.size() == size_t((kGridResolution + 1) * (kGridResolution + 1)));
    nclr::real total_mass = 0;
    for (const auto& c : cells) total_mass += c.mass;
    EXPECT(total_mass > 600.0f && total_mass <= 625.0f + 1e-3f);  // sticky walls only remove mass
    const auto mu = solver->mu_0 * solver->kJellyHardening;
    EXPECT(std::abs(mu - 1000.0f / (2.0f * 1.3f) * 0.3f) < 1e-3f);
    EXPECT(solver->kBoundary == 3 && solver->kSnowHardening == 10.0f && solver->kLiquidHardening == 1.0f);

    // 3D with the reference-mode identity (src/nclr.h:47)
    std::vector<nclr::Particle<3>> p3;
    for (const auto& pos : nclr::cube<3>(8, 0.45, 0.55)) p3.emplace_back(nclr::Particle<3>(pos, 1));
    EXPECT(p3[0].F(2, 2) == 0.0f && p3[0].F(1, 1) == 1.0f);
    nclr::MPMSimulation<3> sim3(p3, nclr::MaterialModel::kSnow, 32);
    sim3.advance(2);
    EXPECT(std::abs(sim3.particles()[0].Jp - 0.6f) < 1e-6f);

    // out-of-range particle: std::out_of_range, as std::vector::at in the reference
    std::vector<nclr::Particle<2>> bad{nclr::Particle<2>(nclr::Vector<nclr::real, 2>(0.5f, 0.5f), 0),
                                       nclr::Particle<2>(nclr::Vector<nclr::real, 2>(1.3f, 0.5f), 0)};
    nclr::MPMSimulation<2> bad_sim(bad, nclr::MaterialModel::kJelly);
    bad_sim.advance();
    bool threw = false;
    try {
        (void)bad_sim.particles();
    } catch (const std::out_of_range&) {
        threw = true;
    }
    EXPECT(threw);

    // host SVD / polar helpers
    nclr::Matrix<nclr::real, 3> A(1.0f, 0.2f, 0.0f, -0.1f, 0.9f, 0.3f, 0.05f, 0.0f, -1.1f), U, S, V, R, P;
    nclr::nclr_svd<3>(A, U, S, V);
    EXPECT((U * S * V.transpose() - A).norm() < 1e-5f);
    EXPECT(std::abs(U.determinant() - 1.0f) < 1e-5f && S(2, 2) < 0.0f);
    nclr::nclr_polar<3>(A, R, P);
    EXPECT((R * P - A).norm() < 1e-5f);
    std::cout << "dropin_example ok: " << steps << " steps, " << ps.size() << " particles" << std::endl;
    return 0;
}
