// Writes a synthetic dump with include/nclr/nclr_dump.h (no GPU): tests/test_dump_io.py loads it
// back with nuclearmpm_amd.ioutils and checks every value, and checks the text formatting.
//   dump_writer <dir> text|npy <dim> <n> <res> <steps>
//   dump_writer format      (prints eigen_text of fixed matrices, separated by "|\n")
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "nclr_dump.h"

// The value of element i of field f at step s; tests/test_dump_io.py computes the same numbers.
static float value(int s, int f, long i) {
    const long h = (i * 2654435761L + f * 97L + s * 7919L) % 20011L;
    const double mag = std::pow(10.0, double(h % 9) - 4.0);
    return float((double(h) / 20011.0 - 0.5) * mag);
}

static std::vector<float> field(int s, int f, size_t count) {
    std::vector<float> a(count);
    for (size_t i = 0; i < count; ++i) a[i] = value(s, f, long(i));
    return a;
}

int main(int argc, char** argv) {
    if (argc == 2 && std::string(argv[1]) == "format") {
        const float v2[2] = {0.5f, -0.25f};
        const float m2[4] = {1.5f, -2.0f, 1e-7f, 100000.0f};
        const float big[3] = {1234567.0f, 0.1f, -3.0f};
        const float lame[2] = {384.615f, 576.923f};
        std::cout << nclr::dump::eigen_text(v2, 2, 1) << "|\n"
                  << nclr::dump::eigen_text(m2, 2, 2) << "|\n"
                  << nclr::dump::eigen_text(big, 3, 1) << "|\n"
                  << nclr::dump::eigen_text(lame, 1, 2) << "|\n";
        return 0;
    }
    if (argc != 7) {
        std::cerr << "usage: dump_writer <dir> text|npy <dim> <n> <res> <steps>\n";
        return 2;
    }
    const std::string dir = argv[1], fmt = argv[2];
    const int d = std::atoi(argv[3]), n = std::atoi(argv[4]), res = std::atoi(argv[5]), steps = std::atoi(argv[6]);
    long nodes = 1;
    for (int a = 0; a < d; ++a) nodes *= res + 1;
    const float dt = 1e-4f, mu = 384.615f * 0.3f, lambda = 576.923f * 0.3f;
    for (int s = 0; s < steps; ++s) {
        const auto x = field(s, 0, size_t(n) * d), v = field(s, 1, size_t(n) * d);
        const auto F = field(s, 2, size_t(n) * d * d), C = field(s, 3, size_t(n) * d * d);
        const auto Jp = field(s, 4, size_t(n)), mass = field(s, 5, size_t(nodes));
        const auto vel = field(s, 6, size_t(nodes) * d);
        if (fmt == "npy") {
            const nclr::dump::NpyDump out(dir);
            out.particles(s, d, n, dt, x.data(), v.data(), F.data(), C.data(), Jp.data(), mu, lambda, res);
            out.grid(s, d, nodes, mass.data(), vel.data());
        } else {
            const nclr::dump::TextDump out(dir);
            out.particles(s, d, n, dt, x.data(), v.data(), F.data(), C.data(), Jp.data(), mu, lambda);
            out.grid(s, d, nodes, mass.data(), vel.data());
        }
    }
    return 0;
}
