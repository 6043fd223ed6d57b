// Per-step state dumps of the headless solver (nuclearmpm_amd/solver/nclr_solver.cpp).
//
// Text format: the files src/solver.cpp:66-131 writes and python/ioutils.py:32-97 reads, one
// file per step and field under the output folder, appended to like the reference does:
//   {step}_timestep.txt  one line per particle: dt * step in float (dt for step 0,
//                        src/solver.cpp:92)
//   {step}_x.txt, _v.txt one coefficient per line (a column vector printed by Eigen)
//   {step}_F.txt, _C.txt one matrix row per line
//   {step}_Jp.txt        one value per line
//   {step}_lame.txt      "mu lambda" per particle: mu_0 and lambda_0 times the material's
//                        hardening constant (src/solver.cpp:100-105)
//   {step}_mass.txt      one value per grid node, x-major (the reference indexes the grid with a
//                        row stride of res instead of res + 1, src/solver.cpp:117,124; here every
//                        node is written once, in order)
//   {step}_velocity.txt  one coefficient per line per node
// Coefficients are formatted the way Eigen's default IOFormat prints them: the stream's default
// precision (6 significant digits), every coefficient right-aligned to the widest one of its
// matrix, separated by one space.
//
// Binary format (--dump-format npy): {step}_{field}.npy, NumPy .npy v1.0 files (little-endian
// float32, C order): x (n,d), v (n,d), F (n,d,d), C (n,d,d), Jp (n,), mass (nodes,),
// velocity (nodes,d), and {step}_meta.npy (float64: step, time, dim, res, n, mu, lambda).
#pragma once

#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace nclr::dump {

// One coefficient as `std::ostream << float` prints it with default flags and precision 6.
inline std::string coeff_text(float v) {
    std::ostringstream s;
    s << v;
    return s.str();
}

// rows x cols row-major values printed like Eigen's operator<< with the default IOFormat.
inline std::string eigen_text(const float* a, int rows, int cols) {
    std::vector<std::string> txt(static_cast<size_t>(rows) * static_cast<size_t>(cols));
    size_t width = 0;
    for (size_t i = 0; i < txt.size(); ++i) {
        txt[i] = coeff_text(a[i]);
        width = std::max(width, txt[i].size());
    }
    std::string out;
    for (int r = 0; r < rows; ++r) {
        if (r) out += '\n';
        for (int c = 0; c < cols; ++c) {
            if (c) out += ' ';
            const std::string& t = txt[static_cast<size_t>(r) * cols + c];
            out.append(width - t.size(), ' ');
            out += t;
        }
    }
    return out;
}

class TextDump {
   public:
    explicit TextDump(std::filesystem::path dir) : dir_(std::move(dir)) { std::filesystem::create_directories(dir_); }

    // Appends `block` + '\n' for each entry, to {step}_{field}.txt.
    void append(int64_t step, const std::string& field, const std::vector<std::string>& blocks) const {
        std::ofstream f(dir_ / (std::to_string(step) + "_" + field + ".txt"), std::ios::app);
        for (const auto& b : blocks) f << b << '\n';
    }

    // Particle fields of one step: x, v (n*d), F, C (n*d*d, row-major), Jp (n).
    void particles(int64_t step, int d, int64_t n, float dt, const float* x, const float* v, const float* F,
                   const float* C, const float* Jp, float mu, float lambda) const {
        const size_t nn = static_cast<size_t>(n);
        std::vector<std::string> ts(nn), xs(nn), vs(nn), Fs(nn), Cs(nn), Js(nn), ls(nn);
        std::ostringstream t;
        t << (step > 0 ? dt * static_cast<float>(step) : dt);
        const float lame[2] = {mu, lambda};
        const std::string lame_txt = eigen_text(lame, 1, 2);
        for (size_t i = 0; i < nn; ++i) {
            ts[i] = t.str();
            xs[i] = eigen_text(x + i * d, d, 1);
            vs[i] = eigen_text(v + i * d, d, 1);
            Fs[i] = eigen_text(F + i * d * d, d, d);
            Cs[i] = eigen_text(C + i * d * d, d, d);
            Js[i] = coeff_text(Jp[i]);
            ls[i] = lame_txt;
        }
        append(step, "timestep", ts);
        append(step, "x", xs);
        append(step, "v", vs);
        append(step, "F", Fs);
        append(step, "C", Cs);
        append(step, "Jp", Js);
        append(step, "lame", ls);
    }

    // Grid of one step: nodes x-major, velocity nodes*d.
    void grid(int64_t step, int d, int64_t nodes, const float* mass, const float* velocity) const {
        std::vector<std::string> ms(static_cast<size_t>(nodes)), vs(static_cast<size_t>(nodes));
        for (int64_t i = 0; i < nodes; ++i) {
            ms[size_t(i)] = coeff_text(mass[i]);
            vs[size_t(i)] = eigen_text(velocity + i * d, d, 1);
        }
        append(step, "mass", ms);
        append(step, "velocity", vs);
    }

   private:
    std::filesystem::path dir_;
};

// NumPy .npy v1.0 writer (float32 or float64, C order).
inline void write_npy(const std::filesystem::path& path, const void* data, const char* descr, size_t elem,
                      const std::vector<int64_t>& shape) {
    std::string dict = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (";
    size_t count = 1;
    for (size_t i = 0; i < shape.size(); ++i) {
        dict += std::to_string(shape[i]);
        dict += (shape.size() == 1 || i + 1 < shape.size()) ? "," : "";
        if (i + 1 < shape.size()) dict += " ";
        count *= static_cast<size_t>(shape[i]);
    }
    dict += "), }";
    const size_t prefix = 10;  // magic (6) + version (2) + header length (2)
    size_t total = prefix + dict.size() + 1;
    const size_t pad = (64 - total % 64) % 64;
    dict.append(pad, ' ');
    dict += '\n';
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    f.write(reinterpret_cast<const char*>(magic), 8);
    const uint16_t hl = static_cast<uint16_t>(dict.size());
    const unsigned char hlb[2] = {static_cast<unsigned char>(hl & 0xff), static_cast<unsigned char>(hl >> 8)};
    f.write(reinterpret_cast<const char*>(hlb), 2);
    f.write(dict.data(), static_cast<std::streamsize>(dict.size()));
    f.write(static_cast<const char*>(data), static_cast<std::streamsize>(count * elem));
}

class NpyDump {
   public:
    explicit NpyDump(std::filesystem::path dir) : dir_(std::move(dir)) { std::filesystem::create_directories(dir_); }

    void particles(int64_t step, int d, int64_t n, float dt, const float* x, const float* v, const float* F,
                   const float* C, const float* Jp, float mu, float lambda, int res) const {
        const auto p = [&](const char* f) { return dir_ / (std::to_string(step) + "_" + f + ".npy"); };
        write_npy(p("x"), x, "<f4", 4, {n, d});
        write_npy(p("v"), v, "<f4", 4, {n, d});
        write_npy(p("F"), F, "<f4", 4, {n, d, d});
        write_npy(p("C"), C, "<f4", 4, {n, d, d});
        write_npy(p("Jp"), Jp, "<f4", 4, {n});
        const float t = step > 0 ? dt * static_cast<float>(step) : dt;
        const double meta[7] = {double(step), double(t), double(d), double(res), double(n), double(mu), double(lambda)};
        write_npy(p("meta"), meta, "<f8", 8, {7});
    }

    void grid(int64_t step, int d, int64_t nodes, const float* mass, const float* velocity) const {
        const auto p = [&](const char* f) { return dir_ / (std::to_string(step) + "_" + f + ".npy"); };
        write_npy(p("mass"), mass, "<f4", 4, {nodes});
        write_npy(p("velocity"), velocity, "<f4", 4, {nodes, d});
    }

   private:
    std::filesystem::path dir_;
};

}  // namespace nclr::dump
