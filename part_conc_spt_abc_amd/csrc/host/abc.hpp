// abc.hpp -- ABC-PMC estimator on the C++ host, calling the hot path through
// the C ABI. Mirrors the reference driver's semantics
// (src/gpu/detail/simulation_impl.inl:178-455, 502-887; src/cpu/estimate.jl:1-240):
// uniform prior on m, log10 c and az; adaptive or fixed initial tolerance;
// per generation gamma -= delta_gamma, each population member re-proposed
// (parent by weight, reflected Gaussian perturbation with tau = sqrt(2 var),
// components sorted by m) until its distance <= gamma; pmc-standard or
// inverse-squared-distance weights; stop when the mean number of trials per
// member reaches ub_average_number_of_trials.
//
// The reference's per-proposal launch queue is replaced by rounds: every
// member still searching in a round is simulated together, as one batch.
#pragma once

#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "../../../include/pcs_abi.h"
#include "io.hpp"

namespace pcs::abc {

// Helpers with closed forms (unit tested).
double var_1n(const std::vector<double>& x);                     // 1/N variance (simulation_impl.inl:24-43)
double normpdf(double x, double mu, double sigma);               // (simulation_impl.inl:46-51)
double reflect_into(double x, double lo, double hi);             // reflection loop of displace_ (:64-79)
std::size_t pick_by_weight(const std::vector<double>& cum, double u);  // lower_bound (:54-62)
double median_upper(std::vector<double> v);                      // nth_element at N/2 (:819-829)
// pmc-standard importance weights (simulation_impl.inl:832-871); rows are
// parameters (m_1..m_C, log10c_1..log10c_C, az_1..az_Z), columns members.
std::vector<double> pmc_weights(const std::vector<std::vector<double>>& theta_new,
                                const std::vector<std::vector<double>>& theta_old, const std::vector<double>& w_old,
                                const std::vector<double>& tau);
std::vector<double> inverse_square_weights(const std::vector<double>& dist);  // (:872-887)

struct Config {
    int device = 0;
    int mode = PCS_MODE_BATCHED;
    std::size_t maxIter = 0;  // 0 = until converged
    int verbosity = 0;
    uint64_t seed = 0;
    // pmc-standard weights move to the GPU (pcs_pmc_weights) from this
    // population size on; below it the host loop is cheaper than the copies.
    std::size_t device_weights_min = 256;
    // Rounds of at least this many trials draw their proposals on the GPU
    // (pcs_batch_propose: parents, displacements, Poisson counts and the
    // batch plan); smaller rounds on the host. 0 = always on the host.
    std::size_t device_propose_min = 2048;
    // The batched walk's launch shape is autotuned (pcs_autotune) once, on
    // the first round of at least this many trials. 0 = never.
    std::size_t autotune_min = 4096;
};

class Estimator {
public:
    Estimator(const io::Parameters& p, const Config& cfg);
    ~Estimator();
    Estimator(const Estimator&) = delete;
    Estimator& operator=(const Estimator&) = delete;

    void run();
    io::Output output() const;

    std::size_t iterations() const { return iterations_; }
    std::size_t simulations() const { return simulations_; }

private:
    struct Trial {
        std::vector<double> m, log10c, az;  // proposed parameters
    };

    // Simulate trials[i] for the given members; writes log10 distances.
    void simulate(const std::vector<Trial>& trials, std::vector<double>& dist);
    // Propose trials.size() trials from the population on the GPU and
    // simulate them; fills trials and dist.
    void simulate_proposed(std::vector<Trial>& trials, std::vector<double>& dist);
    // pcs_autotune on the first calibration proposals of a large round, once.
    void autotune_once(const std::vector<pcs_proposal>& props, const std::vector<uint32_t>& frames,
                       const std::vector<uint32_t>& parts, std::size_t J);
    Trial perturb(std::size_t parent);
    void compute_tau();

    io::Parameters par_;
    Config cfg_;
    std::mt19937_64 rng_;
    std::size_t C_, Z_, N_;
    // population: component-major, [c][member]
    std::vector<std::vector<double>> m_, log10c_, az_;
    std::vector<double> dist_, w_;
    std::vector<double> tauM_, tauC_, tauAz_;
    double gamma_ = 0.0;
    bool converged_ = false;
    std::size_t iterations_ = 0, simulations_ = 0;
    // seconds spent per generation: proposal packing on the host, upload +
    // plan (pcs_batch_load), GPU run + read-back, weights
    double tPack_ = 0.0, tLoad_ = 0.0, tGpu_ = 0.0, tWeights_ = 0.0;
    double lowerC_, upperC_;
    pcs_ctx* ctx_ = nullptr;
    pcs_batch* batch_ = nullptr;
    bool tuned_ = false;
};

}  // namespace pcs::abc
