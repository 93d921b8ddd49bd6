// ref_xml_main.cpp -- driver around the reference's own XML loader and
// writer (src/gpu/shared/input.cpp, output.cpp, common.cpp), built from the
// reference sources by `make ref` into oracle/_ref/. Test infrastructure:
//   ref_xml load <input.xml> <data.xml>   prints the parsed input::Parameters
//   ref_xml write <output.xml>            writes a fixed output::Output
#include <cstdio>
#include <cstring>

#include "shared/input.hpp"
#include "shared/output.hpp"

static void print_list_u(const char* name, const std::vector<std::size_t>& v)
{
    std::printf("\"%s\": [", name);
    for (std::size_t i = 0; i < v.size(); ++i) std::printf("%s%zu", i ? ", " : "", v[i]);
    std::printf("],\n");
}

int main(int argc, char** argv)
{
    if (argc >= 4 && 0 == std::strcmp(argv[1], "load")) {
        input::Parameters p = input::load(argv[2], argv[3]);
        std::printf("{\n");
        std::printf("\"model\": \"%s\",\n", to_string(p.model).c_str());
        std::printf("\"weighting_scheme\": \"%s\",\n", to_string(p.weightingScheme).c_str());
        std::printf("\"number_of_components\": %zu,\n\"kmin\": %zu,\n\"number_of_de_bins\": %zu,\n",
                    p.componentCount, p.kmin, p.deBinCount);
        std::printf("\"number_of_abc_samples\": %zu,\n\"ub_average_number_of_trials\": %zu,\n", p.abcSampleCount,
                    p.avgTrialCount);
        std::printf("\"Lx\": %.17g,\n\"Ly\": %.17g,\n\"Lz\": %.17g,\n", p.Lx, p.Ly, p.Lz);
        std::printf("\"gamma_initial\": %.17g,\n\"delta_gamma\": %.17g,\n\"gamma_adaptive\": %s,\n", p.gamma,
                    p.deltaGamma, p.adaptiveGamma ? "true" : "false");
        std::printf("\"ub_de\": %.17g,\n\"lb_m\": %.17g,\n\"ub_m\": %.17g,\n\"lb_c\": %.17g,\n\"ub_c\": %.17g,\n",
                    p.de.upper, p.m.lower, p.m.upper, p.c.lower, p.c.upper);
        std::printf("\"lb_az\": %.17g,\n\"ub_az\": %.17g,\n", p.az.lower, p.az.upper);
        std::printf("\"output_file_path\": \"%s\",\n", p.outputFilePath.c_str());
        std::printf("\"ax\": %.17g,\n\"ay\": %.17g,\n\"deltat\": %.17g,\n", p.ax, p.ay, p.deltaT);
        print_list_u("number_of_frames", p.frameCounts);
        print_list_u("K", p.Ks);
        std::printf("\"DE\": [");
        for (std::size_t i = 0; i < p.DEs.size(); ++i) std::printf("%s%.17g", i ? ", " : "", p.DEs[i]);
        std::printf("]\n}\n");
        return 0;
    }
    if (argc >= 3 && 0 == std::strcmp(argv[1], "write")) {
        output::Output o;
        o.model = EModel::discreteVariableZ;
        o.weightingScheme = EWeightingScheme::pmcStandard;
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
        output::write(argv[2], o);
        return 0;
    }
    std::fprintf(stderr, "usage: ref_xml load <input.xml> <data.xml> | write <output.xml>\n");
    return 2;
}
