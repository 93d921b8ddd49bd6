// host_tests.cpp -- unit checks of the host side with closed-form answers:
// the ABC-PMC helpers, the XML reader and the XML contract round trip.
// Usage: pcs_host_tests <input.xml> <data.xml> <scratch-dir>; no GPU needed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "abc.hpp"
#include "io.hpp"
#include "xml.hpp"

namespace {

int g_fail = 0, g_checks = 0;

void expect(bool ok, const char* what, int line)
{
    ++g_checks;
    if (!ok) {
        ++g_fail;
        std::fprintf(stderr, "FAIL line %d: %s\n", line, what);
    }
}
#define EXPECT(c) expect((c), #c, __LINE__)

bool near(double a, double b, double tol) { return std::fabs(a - b) <= tol * std::max(1.0, std::fabs(b)); }

template <typename F>
bool throws(F f)
{
    try {
        f();
    } catch (const std::exception&) {
        return true;
    }
    return false;
}

void test_helpers()
{
    using namespace pcs::abc;
    EXPECT(near(var_1n({1, 2, 3, 4}), 1.25, 1e-15));
    EXPECT(var_1n({7}) == 0.0);
    EXPECT(near(normpdf(0, 0, 1), 1.0 / std::sqrt(2 * M_PI), 1e-15));
    EXPECT(near(normpdf(1.3, 0.3, 2.0), normpdf(-0.7, 0.3, 2.0), 1e-15));
    double area = 0.0;
    for (int i = -8000; i < 8000; ++i) area += normpdf(i * 1e-3, 0.0, 0.7) * 1e-3;
    EXPECT(near(area, 1.0, 1e-9));
    EXPECT(near(reflect_into(1.2, 0, 1), 0.8, 1e-15));
    EXPECT(near(reflect_into(-0.3, 0, 1), 0.3, 1e-15));
    EXPECT(near(reflect_into(2.5, 0, 1), 0.5, 1e-15));
    EXPECT(reflect_into(1.0, 0, 1) == 1.0 && reflect_into(0.0, 0, 1) == 0.0);
    const std::vector<double> cum = {0.25, 0.5, 1.0};
    EXPECT(pick_by_weight(cum, 0.1) == 0 && pick_by_weight(cum, 0.25) == 0);
    EXPECT(pick_by_weight(cum, 0.3) == 1 && pick_by_weight(cum, 0.99) == 2);
    EXPECT(pick_by_weight(cum, 1.0) == 2 && pick_by_weight(cum, 1.5) == 2);
    EXPECT(median_upper({5, 1, 3, 2}) == 3.0 && median_upper({4, 9, 1}) == 4.0);
    EXPECT(throws([] { median_upper({}); }));
    // one old particle at 0 with tau 1: weights proportional to 1/phi(theta)
    const auto w = pmc_weights({{0.0, 1.0}}, {{0.0}}, {1.0}, {1.0});
    const double a = 1.0 / normpdf(0, 0, 1), b = 1.0 / normpdf(1, 0, 1);
    EXPECT(near(w[0], a / (a + b), 1e-14) && near(w[1], b / (a + b), 1e-14));
    // two parameters multiply; old weights enter linearly
    const auto w2 = pmc_weights({{0.5}, {-1.0}}, {{0.0, 1.0}, {0.0, 0.0}}, {0.25, 0.75}, {1.0, 2.0});
    EXPECT(near(w2[0], 1.0, 1e-15));
    const auto wi = inverse_square_weights({1.0, 2.0});
    EXPECT(near(wi[0], 0.8, 1e-15) && near(wi[1], 0.2, 1e-15));
}

void test_xml()
{
    using namespace pcs::xml;
    const auto d = parse("<?xml version=\"1.0\"?>\n<!-- c --><root a=\"1\" b='x&amp;y'>"
                         "<k>1 &lt; 2</k><!-- inner --><e/><c><![CDATA[<raw>]]></c><n>&#65;&#x42;</n></root>\n");
    EXPECT(d.root->name == "root" && d.root->attributes.size() == 2 && d.root->attributes[1].second == "x&y");
    EXPECT(d.root->child("k")->text == "1 < 2");
    EXPECT(d.root->child("e") && d.root->child("e")->text.empty());
    EXPECT(d.root->child("c")->text == "<raw>");
    EXPECT(d.root->child("n")->text == "AB");
    EXPECT(d.root->child("missing") == nullptr);
    EXPECT(throws([] { parse("<a><b></a></b>"); }));
    EXPECT(throws([] { parse("<a>"); }));
    EXPECT(throws([] { parse("<a></a><b/>"); }));
    EXPECT(throws([] { parse("<a>&bogus;</a>"); }));
}

void test_contract(const std::string& input, const std::string& data, const std::string& dir)
{
    using namespace pcs::io;
    const Parameters p = load(input, data);
    EXPECT(p.model == Model::discreteVariableZ && p.weightingScheme == WeightingScheme::pmcStandard);
    EXPECT(p.componentCount == 2 && p.kmin == 3 && p.deBinCount == 1000 && p.abcSampleCount == 1000);
    EXPECT(p.frameCounts.size() == 25 && p.Ks.size() == 38302 && p.DEs.size() == 38302);
    EXPECT(p.adaptiveGamma && p.de.upper == 5.0 && p.c.lower == 5e8 && p.deltaT == 0.068);
    // missing data file and unknown enum strings are load errors
    EXPECT(throws([&] { load(input, dir + "/does-not-exist.xml"); }));
    {
        const std::string bad = dir + "/bad_input.xml";
        FILE* f = std::fopen(bad.c_str(), "w");
        std::fprintf(f, "<input><model>lognormal</model></input>");
        std::fclose(f);
        EXPECT(throws([&] { load(bad, data); }));
    }
    EXPECT(throws([] { model_from_string("nope"); }));
    EXPECT(model_from_string("discrete") == Model::discreteFixedZ);
    // output round trip
    Output o;
    o.model = Model::discreteFixedZ;
    o.weightingScheme = WeightingScheme::inverseDistSq;
    o.abcSampleCount = 3;
    o.componentCount = 1;
    o.zComponentCount = 1;
    o.m = {0.1, 0.2, 1.0 / 3.0};
    o.c = {1e8, 2.5e9, 3e9};
    o.az = {1, 2, 3};
    o.dist = {4.5, 5.25, -1.0 / 7.0};
    o.w = {0.2, 0.3, 0.5};
    o.gamma = 3.0625;
    o.converged = false;
    o.meta["k"] = "v";
    o.t_exec = 1.5;
    o.number_of_iterations = 7;
    o.number_of_simulations = 1234;
    const std::string path = dir + "/roundtrip.xml";
    write(path, o);
    const Output r = load_output(path);
    EXPECT(r.m == o.m && r.c == o.c && r.az == o.az && r.dist == o.dist && r.w == o.w);
    EXPECT(r.gamma == o.gamma && !r.converged && r.meta.at("k") == "v");
    EXPECT(r.number_of_simulations == 1234 && r.number_of_iterations == 7 && r.t_exec == 1.5);
    EXPECT(r.model == o.model && r.weightingScheme == o.weightingScheme);
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s <input.xml> <data.xml> <scratch-dir>\n", argv[0]);
        return 2;
    }
    test_helpers();
    test_xml();
    test_contract(argv[1], argv[2], argv[3]);
    std::printf("%d checks, %d failed\n", g_checks, g_fail);
    return g_fail ? 1 : 0;
}
