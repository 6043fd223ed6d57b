// Headless NuclearMPM solver on the MI355X engine: the command line of src/solver.cpp:25-217,
// written against the drop-in header (include/nclr/nclr.h) so that it is also an example of a
// reference caller running unchanged on the engine.
//
// Same flags and defaults as the reference (--steps 1000, --cubes 1, --cube-res 25, --dim 2,
// --E 1000, --nu 0.3, --gravity -100, --material-model jelly, --cube<n>-x / --cube<n>-y, --dump,
// --help; grid resolution 64 and dt 1e-4), the same dump files (include/nclr/nclr_dump.h), and:
//   * --cube<n>-x / --cube<n>-y are passed to cube<dim>(res, min, max) as its min and max
//     (src/solver.cpp:138-144), i.e. every cube is the lattice [x, y]^dim; --cube<n>-min and
//     --cube<n>-max are accepted as clearer spellings of the same two values;
//   * --dim 3 runs (the reference's 3D branch builds an empty simulation, src/solver.cpp:211-216);
//   * the grid dump writes every node once (the reference's row stride is res instead of res + 1,
//     src/solver.cpp:117,124);
//   * dumps are written step by step instead of being held in memory until the end;
//   * extra flags: --res, --dt, --out <dir> (default ./tmp), --dump-format text|npy,
//     --identity-mode 0|1 (mpm_options.identity_mode).
#include <cstdlib>
#include <iostream>
#include <map>
#include <optional>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/nclr/nclr.h"
#include "../../include/nclr/nclr_dump.h"

namespace {

constexpr int kColor = 0xED553B;

void help_msg() {
    std::cout << "Usage: ./nuclear_mpm_solver [OPTIONS]\n"
                 "\tNuclearMPM headless solver (MI355X engine)\n"
                 "Options:\n"
                 "\t--steps\tINTEGER\t[default:1000]\tThe number of simulation steps\n"
                 "\t--cubes\tINTEGER\t[default:1]\tThe number of cubes to add\n"
                 "\t--cube-res\tINTEGER\t[default:25]\tThe resolution of each cube\n"
                 "\t--dim\tINTEGER\t[default:2]\tThe number of dimensions to run the sim in [2 or 3]\n"
                 "\t--E\tFLOAT\t[default:1000.0]\tThe young's modulus of the shape(s)\n"
                 "\t--nu\tFLOAT\t[default:0.3]\tThe poisson's ratio of the shape(s)\n"
                 "\t--gravity\tFLOAT\t[default:-100]\tThe gravitational acceleration on axis 1\n"
                 "\t--material-model\t[jelly, snow, liquid]\t[default:jelly]\tThe material model to use\n"
                 "\t--cube[n]-x, --cube[n]-y\tFLOAT\tCube n spans [x, y] on every axis (also --cube[n]-min/-max)\n"
                 "\t--dump\tDump the particle state and grid at each step\n"
                 "\t--dump-format\t[text, npy]\t[default:text]\tsrc/solver.cpp's text files or NumPy .npy\n"
                 "\t--out\tPATH\t[default:./tmp]\tDump folder\n"
                 "\t--res\tINTEGER\t[default:64]\tGrid resolution\n"
                 "\t--dt\tFLOAT\t[default:1e-4]\tTime step\n"
                 "\t--identity-mode\t[0, 1]\t[default:0]\t1: physical identity instead of diag<dim>(1)\n"
                 "\t--help\tShow this message and exit\n";
}

// --key value, --key=value and bare --flag (value "true").
std::map<std::string, std::string> parse_args(int argc, char** argv) {
    std::map<std::string, std::string> out;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a.rfind("--", 0) != 0) continue;
        a = a.substr(2);
        const auto eq = a.find('=');
        if (eq != std::string::npos) {
            out[a.substr(0, eq)] = a.substr(eq + 1);
        } else if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) {
            out[a] = argv[++i];
        } else {
            out[a] = "true";
        }
    }
    return out;
}

struct Args {
    std::map<std::string, std::string> kv;
    std::optional<std::string> get(const std::string& k) const {
        const auto it = kv.find(k);
        return it == kv.end() ? std::nullopt : std::optional<std::string>(it->second);
    }
    int get_int(const std::string& k, int d) const { return get(k) ? std::stoi(*get(k)) : d; }
    double get_real(const std::string& k, double d) const { return get(k) ? std::stod(*get(k)) : d; }
    bool flag(const std::string& k) const { return get(k) && *get(k) != "false" && *get(k) != "0"; }
};

template <int dim>
std::vector<nclr::Particle<dim>> generate_cubes(const Args& args) {
    std::vector<nclr::Particle<dim>> particles;
    const int cubes = args.get_int("cubes", 1), res = args.get_int("cube-res", 25);
    for (int cc = 0; cc < cubes; ++cc) {
        const std::string k = "cube" + std::to_string(cc);
        auto lo = args.get(k + "-x"), hi = args.get(k + "-y");
        if (!lo) lo = args.get(k + "-min");
        if (!hi) hi = args.get(k + "-max");
        if (!lo || !hi) {
            std::cerr << "Cube: " << cc << " is missing coordinates" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        for (const auto& pos : nclr::cube<dim>(res, nclr::real(std::stod(*lo)), nclr::real(std::stod(*hi))))
            particles.emplace_back(pos, kColor);
    }
    // The Particle ctor sets F = diag<dim>(1) as the reference does; the physical identity mode
    // starts from F = I instead.
    if (args.get_int("identity-mode", 0) == 1)
        for (auto& p : particles) p.F.setIdentity();
    return particles;
}

template <int dim, typename Dumper>
void dump_step(const Dumper& out, int64_t step, const nclr::MPMSimulation<dim>& sim, float dt, float mu, float lambda,
               int res, bool grid_valid) {
    const auto& ps = sim.particles();
    const size_t n = ps.size();
    std::vector<float> x(n * dim), v(n * dim), F(n * dim * dim), C(n * dim * dim), Jp(n);
    for (size_t i = 0; i < n; ++i) {
        for (int a = 0; a < dim; ++a) {
            x[i * dim + a] = ps[i].x(a);
            v[i * dim + a] = ps[i].v(a);
            for (int b = 0; b < dim; ++b) {
                F[(i * dim + a) * dim + b] = ps[i].F(a, b);
                C[(i * dim + a) * dim + b] = ps[i].C(a, b);
            }
        }
        Jp[i] = ps[i].Jp;
    }
    if constexpr (std::is_same_v<Dumper, nclr::dump::NpyDump>)
        out.particles(step, dim, int64_t(n), dt, x.data(), v.data(), F.data(), C.data(), Jp.data(), mu, lambda, res);
    else
        out.particles(step, dim, int64_t(n), dt, x.data(), v.data(), F.data(), C.data(), Jp.data(), mu, lambda);
    int64_t nodes = 1;
    for (int a = 0; a < dim; ++a) nodes *= res + 1;
    std::vector<float> mass(size_t(nodes), 0.0f), vel(size_t(nodes) * dim, 0.0f);
    if (grid_valid) {  // the grid after the previous step's update; an empty grid before step 1
        const auto& cells = sim.grid();
        for (int64_t i = 0; i < nodes; ++i) {
            mass[size_t(i)] = cells[size_t(i)].mass;
            for (int a = 0; a < dim; ++a) vel[size_t(i) * dim + a] = cells[size_t(i)].velocity(a);
        }
    }
    out.grid(step, dim, nodes, mass.data(), vel.data());
}

template <int dim>
int run(const Args& args, nclr::MaterialModel model, const std::string& material) {
    const int steps = args.get_int("steps", 1000), res = args.get_int("res", 64);
    const float dt = float(args.get_real("dt", 1e-4));
    const bool dump = args.flag("dump");
    const std::string format = args.get("dump-format").value_or("text");
    const std::filesystem::path dir = args.get("out").value_or((std::filesystem::current_path() / "tmp").string());
    mpm_options opts;
    mpm_default_options(&opts);
    opts.identity_mode = args.get_int("identity-mode", 0);
    nclr::MPMSimulation<dim> sim(generate_cubes<dim>(args), model, res, dt,
                                 nclr::real(args.get_real("E", 1000.0)), nclr::real(args.get_real("nu", 0.3)),
                                 nclr::real(args.get_real("gravity", -100.0)), &opts);
    // lame.txt: mu_0 and lambda_0 times the material's hardening constant (src/solver.cpp:100-105)
    const nclr::real e = material == "snow"    ? nclr::MPMSimulation<dim>::kSnowHardening
                         : material == "jelly" ? nclr::MPMSimulation<dim>::kJellyHardening
                                               : nclr::MPMSimulation<dim>::kLiquidHardening;
    const float mu = float(sim.mu_0 * e), lambda = float(sim.lambda_0 * e);
    std::cout << "Running simulation" << std::endl;
    if (!dump) {
        sim.advance(steps);
        sim.synchronize();
    } else if (format == "npy") {
        const nclr::dump::NpyDump out(dir);
        for (int step = 0; step < steps; ++step) {
            dump_step<dim>(out, step, sim, dt, mu, lambda, res, step > 0);
            sim.advance();
        }
    } else if (format == "text") {
        const nclr::dump::TextDump out(dir);
        for (int step = 0; step < steps; ++step) {
            dump_step<dim>(out, step, sim, dt, mu, lambda, res, step > 0);
            sim.advance();
        }
    } else {
        std::cerr << "Invalid Option: --dump-format " << format << std::endl;
        return EXIT_FAILURE;
    }
    sim.synchronize();
    std::cout << "Simulation done" << std::endl;
    return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv) {
    const Args args{parse_args(argc, argv)};
    const auto material = args.get("material-model");
    if (material && *material != "jelly" && *material != "snow" && *material != "liquid") {
        std::cerr << "Invalid Option: " << *material << std::endl;
        help_msg();
        return EXIT_FAILURE;
    }
    if (args.flag("help") || args.kv.empty()) {
        help_msg();
        if (args.flag("help")) return EXIT_SUCCESS;
    }
    const std::string m = material.value_or("jelly");
    const auto model = m == "snow" ? nclr::MaterialModel::kSnow
                       : m == "liquid" ? nclr::MaterialModel::kLiquid
                                       : nclr::MaterialModel::kJelly;
    try {
        const int dim = args.get_int("dim", 2);
        if (dim == 2) return run<2>(args, model, m);
        if (dim == 3) return run<3>(args, model, m);
        std::cerr << "--dim must be 2 or 3" << std::endl;
        return EXIT_FAILURE;
    } catch (const std::exception& ex) {
        std::cerr << "error: " << ex.what() << std::endl;
        return EXIT_FAILURE;
    }
}
