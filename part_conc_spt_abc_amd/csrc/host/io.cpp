// io.cpp -- input/data/output XML contract (see io.hpp).
#include "io.hpp"

#include <cctype>
#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "xml.hpp"

namespace pcs::io {

namespace {

std::string fmt(const char* f, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, f);
    std::vsnprintf(buf, sizeof(buf), f, ap);
    va_end(ap);
    return buf;
}

const xml::Element* child_checked(const xml::Element* parent, const char* name)
{
    if (const auto* c = parent->child(name)) return c;
    throw XMLMissing(fmt("No <%s> in <%s>", name, parent->name.c_str()));
}

std::string trimmed(const std::string& s)
{
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}

long long child_int64(const xml::Element* parent, const char* name)
{
    const std::string t = trimmed(child_checked(parent, name)->text);
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(t.c_str(), &end, 10);
    if (t.empty() || *end != '\0' || errno)
        throw XMLGeneric(fmt("Can't parse '%s' as an integer for <%s> in <%s>", t.c_str(), name, parent->name.c_str()));
    return v;
}

std::size_t child_count(const xml::Element* parent, const char* name)
{
    const long long v = child_int64(parent, name);
    if (v < 0) throw XMLGeneric(fmt("<%s> in <%s> must not be negative", name, parent->name.c_str()));
    return (std::size_t)v;
}

double child_double(const xml::Element* parent, const char* name)
{
    const std::string t = trimmed(child_checked(parent, name)->text);
    char* end = nullptr;
    const double v = std::strtod(t.c_str(), &end);
    if (t.empty() || *end != '\0')
        throw XMLGeneric(fmt("Can't parse '%s' as a number for <%s> in <%s>", t.c_str(), name, parent->name.c_str()));
    return v;
}

// true/false, or an integer (non-zero is true), like tinyxml2's QueryBoolText.
bool child_bool(const xml::Element* parent, const char* name)
{
    const std::string t = trimmed(child_checked(parent, name)->text);
    if (t == "true") return true;
    if (t == "false") return false;
    char* end = nullptr;
    const long v = std::strtol(t.c_str(), &end, 10);
    if (!t.empty() && *end == '\0') return v != 0;
    throw XMLGeneric(fmt("Can't parse '%s' as a boolean for <%s> in <%s>", t.c_str(), name, parent->name.c_str()));
}

Bounds bounds(const xml::Element* parent, const char* stem)
{
    Bounds b;
    b.lower = child_double(parent, (std::string("lb_") + stem).c_str());
    b.upper = child_double(parent, (std::string("ub_") + stem).c_str());
    return b;
}

// Comma separated list (src/gpu/shared/input.cpp:133-152); surrounding
// whitespace is tolerated.
template <typename T, typename Parse>
std::vector<T> parse_list(const xml::Element* parent, const char* name, Parse parse)
{
    const std::string t = child_checked(parent, name)->text;
    std::vector<T> out;
    const char* p = t.c_str();
    for (;;) {
        while (std::isspace((unsigned char)*p)) ++p;
        if (!*p) break;
        char* end = nullptr;
        const T v = parse(p, &end);
        if (end == p)
            throw XMLParsing(fmt("Can't parse list <%s> in <%s>. Unparsable mayhem '%.40s' remains.", name,
                                 parent->name.c_str(), p));
        out.push_back(v);
        p = end;
        while (std::isspace((unsigned char)*p)) ++p;
        if (*p == ',') ++p;
        else if (*p)
            throw XMLParsing(fmt("Can't parse list <%s> in <%s>. Unparsable mayhem '%.40s' remains.", name,
                                 parent->name.c_str(), p));
    }
    return out;
}

std::string join(const std::vector<double>& v)
{
    std::string s;
    char buf[64];
    for (size_t i = 0; i < v.size(); ++i) {
        std::snprintf(buf, sizeof(buf), i + 1 < v.size() ? "%.18g, " : "%.18g", v[i]);
        s += buf;
    }
    return s;
}

}  // namespace

std::string to_string(Model m)
{
    return m == Model::discreteFixedZ ? "discrete-fixed-depth" : "discrete-variable-depth";
}

std::string to_string(WeightingScheme w)
{
    return w == WeightingScheme::pmcStandard ? "pmc-standard" : "inverse-distance-squared";
}

// src/gpu/shared/common.cpp:40-74 ("discrete" is the deprecated alias of
// the fixed-depth model; "lognormal" is not supported by either simulator).
Model model_from_string(const std::string& s)
{
    if (s == "discrete-fixed-depth") return Model::discreteFixedZ;
    if (s == "discrete-variable-depth") return Model::discreteVariableZ;
    if (s == "discrete") {
        std::fprintf(stderr, "WARNING: EModel: deprecated value \"discrete\"\n");
        return Model::discreteFixedZ;
    }
    throw EnumUnknownString("'" + s + "' is not a valid EModel");
}

WeightingScheme weighting_from_string(const std::string& s)
{
    if (s == "pmc-standard") return WeightingScheme::pmcStandard;
    if (s == "inverse-distance-squared") return WeightingScheme::inverseDistSq;
    throw EnumUnknownString("'" + s + "' is not a valid EWeightingScheme");
}

Parameters load(const std::string& inputPath, const std::string& dataPathIn)
{
    std::string dataPath = dataPathIn;
    try {
        const xml::Document input = xml::load_file(inputPath);
        const xml::Element* iroot = input.root.get();
        if (iroot->name != "input") throw XMLMissing("No <input> root element (input)");
        if (dataPath.empty()) dataPath = trimmed(child_checked(iroot, "data_file_path")->text);
        const xml::Document data = xml::load_file(dataPath);
        const xml::Element* droot = data.root.get();
        if (droot->name != "data") throw XMLMissing("No <data> root element (data)");

        Parameters p;
        const xml::Element* model = iroot->child("model");
        if (!model) {
            model = iroot->child("distribution_class");
            if (model) std::fprintf(stderr, "WARNING: using deprecated \"distribution_class\" field\n");
        }
        if (!model) throw XMLMissing("No <model> in <input>");
        p.model = model_from_string(trimmed(model->text));
        p.weightingScheme = weighting_from_string(trimmed(child_checked(iroot, "weighting_scheme")->text));
        p.componentCount = child_count(iroot, "number_of_components");
        p.kmin = child_count(iroot, "kmin");
        p.deBinCount = child_count(iroot, "number_of_de_bins");
        p.abcSampleCount = child_count(iroot, "number_of_abc_samples");
        p.avgTrialCount = child_count(iroot, "ub_average_number_of_trials");
        p.Lx = child_double(iroot, "Lx");
        p.Ly = child_double(iroot, "Ly");
        p.Lz = child_double(iroot, "Lz");
        p.gamma = child_double(iroot, "gamma_initial");
        p.deltaGamma = child_double(iroot, "delta_gamma");
        if (iroot->child("gamma_adaptive")) {
            try {
                p.adaptiveGamma = child_bool(iroot, "gamma_adaptive");
            } catch (const XMLLoadError&) {
                std::fprintf(stderr, "Note: <gamma_adaptive> not set/invalid (old input?); defaulting to false.\n");
                p.adaptiveGamma = false;
            }
        } else {
            std::fprintf(stderr, "Note: <gamma_adaptive> not set/invalid (old input?); defaulting to false.\n");
        }
        p.de.lower = 0.0;
        p.de.upper = child_double(iroot, "ub_de");
        p.m = bounds(iroot, "m");
        if (iroot->child("lb_s") && iroot->child("ub_s")) {
            p.s = bounds(iroot, "s");
            std::fprintf(stderr, "Note: deprecated <*_s> tags found (old input?)\n");
        }
        p.c = bounds(iroot, "c");
        p.az = bounds(iroot, "az");
        p.outputFilePath = trimmed(child_checked(iroot, "output_file_path")->text);

        p.ax = child_double(droot, "ax");
        p.ay = child_double(droot, "ay");
        p.deltaT = child_double(droot, "deltat");
        p.frameCounts = parse_list<std::size_t>(droot, "number_of_frames",
                                                [](const char* a, char** b) { return (std::size_t)std::strtoull(a, b, 0); });
        p.Ks = parse_list<std::size_t>(droot, "K", [](const char* a, char** b) { return (std::size_t)std::strtoull(a, b, 0); });
        p.DEs = parse_list<double>(droot, "DE", [](const char* a, char** b) { return std::strtod(a, b); });
        if (p.Ks.size() != p.DEs.size())
            throw XMLGeneric(fmt("<K> has %zu entries but <DE> has %zu", p.Ks.size(), p.DEs.size()));
        return p;
    } catch (const std::exception& e) {
        throw XMLLoadError(fmt("Error loading input=%s and data=%s: %s", inputPath.c_str(),
                               dataPath.empty() ? "(unknown)" : dataPath.c_str(), e.what()));
    }
}

void write(const std::string& path, const Output& o)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write '" + path + "'");
    std::fprintf(f, "<output>\n");
    std::fprintf(f, "\t<model>%s</model>\n", to_string(o.model).c_str());
    std::fprintf(f, "\t<weighting_scheme>%s</weighting_scheme>\n", to_string(o.weightingScheme).c_str());
    std::fprintf(f, "\t<number_of_abc_samples>%zu</number_of_abc_samples>\n", o.abcSampleCount);
    std::fprintf(f, "\t<number_of_components>%zu</number_of_components>\n", o.componentCount);
    std::fprintf(f, "\t<number_of_z_components>%zu</number_of_z_components>\n", o.zComponentCount);
    std::fprintf(f, "\n");
    std::fprintf(f, "\t<meta>\n");
    for (const auto& kv : o.meta) std::fprintf(f, "\t\t<%s>%s</%s>\n", kv.first.c_str(), kv.second.c_str(), kv.first.c_str());
    std::fprintf(f, "\t</meta>\n");
    std::fprintf(f, "\n");
    std::fprintf(f, "\t<converged>%s</converged>\n", o.converged ? "true" : "false");
    std::fprintf(f, "\t<gamma>%.18g</gamma>\n", o.gamma);
    std::fprintf(f, "\t<m>%s\t</m>\n", join(o.m).c_str());
    std::fprintf(f, "\t<c>%s\t</c>\n", join(o.c).c_str());
    std::fprintf(f, "\t<az>%s\t</az>\n", join(o.az).c_str());
    std::fprintf(f, "\t<dist>%s\t</dist>\n", join(o.dist).c_str());
    std::fprintf(f, "\t<w>%s\t</w>\n", join(o.w).c_str());
    if (o.t_exec >= 0.0) std::fprintf(f, "\t<t_exec>%.18g</t_exec>\n", o.t_exec);
    if (o.number_of_iterations >= 0)
        std::fprintf(f, "\t<number_of_iterations>%lld</number_of_iterations>\n", o.number_of_iterations);
    if (o.number_of_simulations >= 0)
        std::fprintf(f, "\t<number_of_simulations>%lld</number_of_simulations>\n", o.number_of_simulations);
    std::fprintf(f, "</output>\n");
    if (std::fclose(f) != 0) throw std::runtime_error("error writing '" + path + "'");
}

Output load_output(const std::string& path)
{
    const xml::Document d = xml::load_file(path);
    const xml::Element* r = d.root.get();
    if (r->name != "output") throw XMLMissing("No <output> root element");
    Output o;
    o.model = model_from_string(trimmed(child_checked(r, "model")->text));
    o.weightingScheme = weighting_from_string(trimmed(child_checked(r, "weighting_scheme")->text));
    o.abcSampleCount = child_count(r, "number_of_abc_samples");
    o.componentCount = child_count(r, "number_of_components");
    o.zComponentCount = r->child("number_of_z_components") ? child_count(r, "number_of_z_components")
                                                           : o.componentCount;
    o.converged = child_bool(r, "converged");
    o.gamma = child_double(r, "gamma");
    auto dbl = [](const char* a, char** b) { return std::strtod(a, b); };
    o.m = parse_list<double>(r, "m", dbl);
    o.c = parse_list<double>(r, "c", dbl);
    o.az = parse_list<double>(r, "az", dbl);
    o.dist = parse_list<double>(r, "dist", dbl);
    o.w = parse_list<double>(r, "w", dbl);
    if (const auto* meta = r->child("meta"))
        for (const auto& kv : meta->children) o.meta[kv->name] = trimmed(kv->text);
    if (r->child("t_exec")) o.t_exec = child_double(r, "t_exec");
    if (r->child("number_of_iterations")) o.number_of_iterations = child_int64(r, "number_of_iterations");
    if (r->child("number_of_simulations")) o.number_of_simulations = child_int64(r, "number_of_simulations");
    return o;
}

}  // namespace pcs::io
