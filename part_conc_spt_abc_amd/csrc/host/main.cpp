// main.cpp -- pcs_abc, the command-line estimator (the reference's accel
// binary, src/gpu/main.cpp:61-317): reads input.xml + data.xml, runs
// ABC-PMC with the MI355X hot path, writes output.xml.
#include <unistd.h>

#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <random>
#include <stdexcept>
#include <string>

#include "../../../include/pcs_abi.h"
#include "abc.hpp"
#include "io.hpp"

namespace {

struct Cli {
    std::string input = "./input.xml", data, output;
    int device = 0;
    int mode = PCS_MODE_BATCHED;
    std::size_t maxIter = 0;
    bool fixedSeed = false;
    uint64_t seed = 0;
    int verbosity = 0;
    bool listDevices = false, parseOnly = false;
    std::string writeSample;
    std::size_t deviceProposeMin = 2048, autotuneMin = 4096;
};

void help(FILE* f, const char* self)
{
    std::fprintf(f,
                 "Synopsis: %s [--flags...]\n\n"
                 "Flags:\n"
                 "  --input, -i FILE         : input XML file (default ./input.xml)\n"
                 "  --data, -d FILE          : override the data XML location given in the input\n"
                 "  --output, -o FILE        : override the output XML destination given in the input\n"
                 "  --device N               : GPU index (default 0)\n"
                 "  --gpus, -g SPEC          : accepted for compatibility; the first listed GPU (1-based) is used\n"
                 "  --mode plain|batched     : one launch per simulation, or one batch per round (default)\n"
                 "  --max-steps, -S N        : stop after N generations\n"
                 "  --fixed-seed, -F N       : seed the host random numbers with N\n"
                 "  --quiet, -q / --verbose, -v\n"
                 "  --device-propose-min N   : draw a round's proposals on the GPU from N trials on (default\n"
                 "                             2048; 0 = always on the host)\n"
                 "  --autotune-min N         : autotune the walk's launch shape on the first round of N or\n"
                 "                             more trials (default 4096; 0 = never)\n"
                 "  --list-devices           : list GPUs and exit\n"
                 "  --parse-only             : print the parsed input as JSON and exit (no GPU)\n"
                 "  --write-sample FILE      : write a fixed output XML and exit (no GPU)\n"
                 "  --help, -h\n",
                 self);
}

Cli parse(int argc, char** argv)
{
    Cli c;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto value = [&](const char* what) -> std::string {
            if (i + 1 >= argc) throw std::runtime_error("Command line argument " + a + " requires " + what);
            return argv[++i];
        };
        auto number = [&](const char* what) -> unsigned long long {
            const std::string v = value(what);
            char* end = nullptr;
            const unsigned long long n = std::strtoull(v.c_str(), &end, 10);
            if (v.empty() || *end || v[0] == '-')
                throw std::runtime_error("Command line argument " + a + " requires " + what + " and not " + v);
            return n;
        };
        if (a == "--help" || a == "-h") {
            help(stdout, argv[0]);
            std::exit(0);
        } else if (a == "--quiet" || a == "-q") --c.verbosity;
        else if (a == "--verbose" || a == "-v") ++c.verbosity;
        else if (a == "--input" || a == "-i") c.input = value("a file");
        else if (a == "--data" || a == "-d") c.data = value("a file");
        else if (a == "--output" || a == "-o") c.output = value("a file");
        else if (a == "--device") c.device = (int)number("a device index");
        else if (a == "--gpus" || a == "-g") {
            const std::string spec = value("a GPU spec");
            unsigned dev = 1;
            if (std::sscanf(spec.c_str(), "%u", &dev) != 1 || dev == 0)
                throw std::runtime_error("Don't understand provided gpuspec '" + spec + "'");
            c.device = (int)dev - 1;
        } else if (a == "--mode") {
            const std::string m = value("plain or batched");
            if (m == "plain") c.mode = PCS_MODE_PLAIN;
            else if (m == "batched") c.mode = PCS_MODE_BATCHED;
            else throw std::runtime_error("--mode must be plain or batched, not " + m);
        } else if (a == "--max-steps" || a == "-S") c.maxIter = number("a positive integer");
        else if (a == "--fixed-seed" || a == "-F") {
            c.seed = number("an integer");
            c.fixedSeed = true;
        } else if (a == "--host-scalar" || a == "-hs" || a == "--dev-scalar" || a == "-ds" || a == "--dist-type" ||
                   a == "-dt") {
            const std::string v = value("f or d");
            if (v != "f" && v != "d")
                throw std::runtime_error("Scalar type must be either 'f' (float) or 'd' (double) and not '" + v + "'");
            std::fprintf(stderr, "Note: %s ignored: host FP64, device FP32, exact integer distance\n", a.c_str());
        } else if (a == "--device-propose-min") c.deviceProposeMin = number("a number of trials");
        else if (a == "--autotune-min") c.autotuneMin = number("a number of trials");
        else if (a == "--list-devices") c.listDevices = true;
        else if (a == "--parse-only") c.parseOnly = true;
        else if (a == "--write-sample") c.writeSample = value("a file");
        else throw std::runtime_error("Unknown command line argument '" + a + "'");
    }
    return c;
}

void print_json(const pcs::io::Parameters& p)
{
    std::printf("{\n\"model\": \"%s\",\n\"weighting_scheme\": \"%s\",\n", pcs::io::to_string(p.model).c_str(),
                pcs::io::to_string(p.weightingScheme).c_str());
    std::printf("\"number_of_components\": %zu,\n\"kmin\": %zu,\n\"number_of_de_bins\": %zu,\n", p.componentCount, p.kmin,
                p.deBinCount);
    std::printf("\"number_of_abc_samples\": %zu,\n\"ub_average_number_of_trials\": %zu,\n", p.abcSampleCount,
                p.avgTrialCount);
    std::printf("\"Lx\": %.17g,\n\"Ly\": %.17g,\n\"Lz\": %.17g,\n", p.Lx, p.Ly, p.Lz);
    std::printf("\"gamma_initial\": %.17g,\n\"delta_gamma\": %.17g,\n\"gamma_adaptive\": %s,\n", p.gamma, p.deltaGamma,
                p.adaptiveGamma ? "true" : "false");
    std::printf("\"ub_de\": %.17g,\n\"lb_m\": %.17g,\n\"ub_m\": %.17g,\n\"lb_c\": %.17g,\n\"ub_c\": %.17g,\n", p.de.upper,
                p.m.lower, p.m.upper, p.c.lower, p.c.upper);
    std::printf("\"lb_az\": %.17g,\n\"ub_az\": %.17g,\n\"output_file_path\": \"%s\",\n", p.az.lower, p.az.upper,
                p.outputFilePath.c_str());
    std::printf("\"ax\": %.17g,\n\"ay\": %.17g,\n\"deltat\": %.17g,\n", p.ax, p.ay, p.deltaT);
    auto list = [](const char* n, auto const& v, const char* f, bool last) {
        std::printf("\"%s\": [", n);
        for (std::size_t i = 0; i < v.size(); ++i) {
            if (i) std::printf(", ");
            std::printf(f, v[i]);
        }
        std::printf(last ? "]\n" : "],\n");
    };
    list("number_of_frames", p.frameCounts, "%zu", false);
    list("K", p.Ks, "%zu", false);
    list("DE", p.DEs, "%.17g", true);
    std::printf("}\n");
}

std::string timestamp(std::chrono::system_clock::time_point t)
{
    const std::time_t tt = std::chrono::system_clock::to_time_t(t);
    char buf[64];
    std::strftime(buf, sizeof(buf), "%F %T%z", std::localtime(&tt));
    return buf;
}

}  // namespace

int main(int argc, char** argv)
try {
    const Cli cli = parse(argc, argv);
    if (cli.listDevices) {
        const int n = pcs_device_count();
        std::printf("%d GPU(s) visible\n", n);
        for (int i = 0; i < n; ++i) std::printf("  %d: use --device %d (or -g %d)\n", i, i, i + 1);
        return 0;
    }
    if (!cli.writeSample.empty()) {
        pcs::io::Output o;
        o.model = pcs::io::Model::discreteVariableZ;
        o.weightingScheme = pcs::io::WeightingScheme::pmcStandard;
        o.abcSampleCount = 2;
        o.componentCount = 2;
        o.zComponentCount = 2;
        o.m = {1.5, 2.5, 0.25, 3.0};
        o.c = {1e9, 2e9, 3.5e8, 4e8};
        o.az = {1.0, 2.0, 1.5, 0.5};
        o.dist = {3.25, 4.5};
        o.w = {0.25, 0.75};
        o.converged = true;
        o.gamma = 2.125;
        o.meta["simulator"] = "gpu";
        pcs::io::write(cli.writeSample, o);
        return 0;
    }
    pcs::io::Parameters par = pcs::io::load(cli.input, cli.data);
    if (!cli.output.empty()) par.outputFilePath = cli.output;
    if (cli.parseOnly) {
        print_json(par);
        return 0;
    }

    pcs::abc::Config cfg;
    cfg.device = cli.device;
    cfg.mode = cli.mode;
    cfg.maxIter = cli.maxIter;
    cfg.verbosity = cli.verbosity;
    cfg.seed = cli.fixedSeed ? cli.seed : std::random_device{}();
    cfg.device_propose_min = cli.deviceProposeMin;
    cfg.autotune_min = cli.autotuneMin;

    const auto start = std::chrono::system_clock::now();
    const auto t0 = std::chrono::steady_clock::now();
    pcs::abc::Estimator est(par, cfg);
    est.run();
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const auto end = std::chrono::system_clock::now();

    pcs::io::Output out = est.output();
    out.meta["runtime_ms"] = std::to_string((long long)(secs * 1e3));
    out.meta["started_at"] = timestamp(start);
    out.meta["finished_at"] = timestamp(end);
    char host[256] = {};
    if (0 == ::gethostname(host, sizeof(host) - 1)) out.meta["machine"] = host;
#ifdef NDEBUG
    out.meta["debug"] = "false";
#else
    out.meta["debug"] = "true";
#endif
    out.meta["input_file"] = cli.input;
    if (!cli.data.empty()) out.meta["data_file"] = cli.data;
    out.meta["seed"] = std::to_string(cfg.seed);
    out.t_exec = secs;
    pcs::io::write(par.outputFilePath, out);
    if (cli.verbosity >= 0)
        std::printf("Output written to %s (%zu generations, %zu simulations, %.2f s)\n", par.outputFilePath.c_str(),
                    est.iterations(), est.simulations(), secs);
    return 0;
} catch (const std::exception& e) {
    std::fprintf(stderr, "Top-level exception: %s\n", e.what());
    return 1;
}
