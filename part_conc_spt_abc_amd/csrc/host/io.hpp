// io.hpp -- the input-XML / data-XML / output-XML contract of the reference
// (src/gpu/shared/input.hpp:19-54, src/gpu/shared/output.hpp:16-36,
// src/gpu/shared/common.hpp:16-39), kept key for key.
#pragma once

#include <cstddef>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace pcs::io {

enum class Model { discreteFixedZ, discreteVariableZ };
enum class WeightingScheme { pmcStandard, inverseDistSq };

std::string to_string(Model);
std::string to_string(WeightingScheme);
Model model_from_string(const std::string&);                  // throws EnumUnknownString
WeightingScheme weighting_from_string(const std::string&);    // throws EnumUnknownString

// Error kinds of src/gpu/shared/errors.hpp:6-41.
struct XMLLoadError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct XMLMissing : XMLLoadError {
    using XMLLoadError::XMLLoadError;
};
struct XMLGeneric : XMLLoadError {
    using XMLLoadError::XMLLoadError;
};
struct XMLParsing : XMLLoadError {
    using XMLLoadError::XMLLoadError;
};
struct EnumUnknownString : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Bounds {
    double lower = 0.0, upper = 0.0;
};

// input::Parameters (src/gpu/shared/input.hpp:19-45).
struct Parameters {
    Model model = Model::discreteFixedZ;
    WeightingScheme weightingScheme = WeightingScheme::pmcStandard;
    std::size_t componentCount = 0, kmin = 0, deBinCount = 0, abcSampleCount = 0, avgTrialCount = 0;
    double Lx = 0, Ly = 0, Lz = 0;
    double gamma = 0, deltaGamma = 0;
    bool adaptiveGamma = false;
    Bounds de, m, s, c, az;
    std::string outputFilePath;
    // data.xml
    double ax = 0, ay = 0, deltaT = 0;
    std::vector<std::size_t> frameCounts;
    std::vector<std::size_t> Ks;
    std::vector<double> DEs;
};

// input::load (src/gpu/shared/input.cpp:158-291): the data file comes from
// <data_file_path> unless dataPath overrides it. Errors are XMLLoadError
// whose message names both files and the cause.
Parameters load(const std::string& inputPath, const std::string& dataPath = "");

// output::Output (src/gpu/shared/output.hpp:16-34). meta keeps insertion
// order per key (std::map: sorted, deterministic).
struct Output {
    Model model = Model::discreteFixedZ;
    WeightingScheme weightingScheme = WeightingScheme::pmcStandard;
    std::size_t abcSampleCount = 0, componentCount = 0, zComponentCount = 0;
    std::vector<double> m, c, az, dist, w;  // m/c/az column-major: component fastest
    bool converged = false;
    double gamma = 0.0;
    std::map<std::string, std::string> meta;
    // Keys of the Julia writer (src/cpu/file_io/write_output.jl:29-32), written
    // after <w> when set (>= 0), so either reader finds what it looks for.
    double t_exec = -1.0;
    long long number_of_iterations = -1, number_of_simulations = -1;
};

// output::write (src/gpu/shared/output.cpp:232-306): same element order and
// number formatting (%.18g, ", " separators).
void write(const std::string& path, const Output& out);

// Reads back what write() produced (the reference's output::load is a stub,
// src/gpu/shared/output.cpp:127-130).
Output load_output(const std::string& path);

}  // namespace pcs::io
