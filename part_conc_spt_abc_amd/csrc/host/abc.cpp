// abc.cpp -- ABC-PMC estimator (see abc.hpp).
#include "abc.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace pcs::abc {

double var_1n(const std::vector<double>& x)
{
    if (x.empty()) return 0.0;
    double mean = 0.0;
    for (double v : x) mean += v;
    mean /= (double)x.size();
    double s = 0.0;
    for (double v : x) s += (v - mean) * (v - mean);
    return s / (double)x.size();
}

double normpdf(double x, double mu, double sigma)
{
    const double xm = x - mu;
    return std::exp(-0.5 * xm * xm / (sigma * sigma)) / (std::sqrt(2.0 * M_PI) * sigma);
}

double reflect_into(double x, double lo, double hi)
{
    while (x < lo || x > hi) {
        if (x < lo) x += 2.0 * (lo - x);
        if (x > hi) x -= 2.0 * (x - hi);
    }
    return x;
}

std::size_t pick_by_weight(const std::vector<double>& cum, double u)
{
    const auto it = std::lower_bound(cum.begin(), cum.end(), u);
    return std::min<std::size_t>(cum.size() - 1, (std::size_t)(it - cum.begin()));
}

double median_upper(std::vector<double> v)
{
    if (v.empty()) throw std::invalid_argument("median of an empty set");
    const std::size_t n = v.size() / 2;
    std::nth_element(v.begin(), v.begin() + (std::ptrdiff_t)n, v.end());
    return v[n];
}

std::vector<double> pmc_weights(const std::vector<std::vector<double>>& theta_new,
                                const std::vector<std::vector<double>>& theta_old, const std::vector<double>& w_old,
                                const std::vector<double>& tau)
{
    const std::size_t P = theta_new.size();
    const std::size_t N = P ? theta_new[0].size() : 0;
    const std::size_t Nold = w_old.size();
    std::vector<double> w(N, 0.0);
    double wsum = 0.0;
    for (std::size_t j = 0; j < N; ++j) {
        double acc = 0.0;
        for (std::size_t i = 0; i < Nold; ++i) {
            double term = 1.0;
            for (std::size_t p = 0; p < P; ++p) term *= normpdf(theta_new[p][j] - theta_old[p][i], 0.0, tau[p]);
            acc += w_old[i] * term;
        }
        w[j] = 1.0 / acc;
        wsum += w[j];
    }
    for (double& v : w) v /= wsum;
    return w;
}

std::vector<double> inverse_square_weights(const std::vector<double>& dist)
{
    std::vector<double> w(dist.size());
    double s = 0.0;
    for (std::size_t i = 0; i < dist.size(); ++i) {
        w[i] = 1.0 / (dist[i] * dist[i]);
        s += w[i];
    }
    for (double& v : w) v /= s;
    return w;
}

namespace {

void check(int r, const char* what)
{
    if (r != PCS_OK) throw std::runtime_error(std::string(what) + ": " + pcs_last_error());
}

}  // namespace

Estimator::Estimator(const io::Parameters& p, const Config& cfg)
    : par_(p), cfg_(cfg), rng_(cfg.seed)
{
    C_ = p.componentCount;
    Z_ = p.model == io::Model::discreteFixedZ ? 1 : C_;
    N_ = p.abcSampleCount;
    if (C_ < 1 || C_ > PCS_MAX_SPECIES) throw std::invalid_argument("number_of_components must be in [1, 8]");
    if (N_ < 1) throw std::invalid_argument("number_of_abc_samples must be positive");
    if (p.frameCounts.empty()) throw std::invalid_argument("data has no videos");
    if (p.deBinCount < 1 || p.deBinCount > PCS_MAX_NDE) throw std::invalid_argument("number_of_de_bins out of range");
    if (!(p.c.lower > 0.0) || !(p.c.upper >= p.c.lower)) throw std::invalid_argument("concentration bounds invalid");
    lowerC_ = std::log10(p.c.lower);
    upperC_ = std::log10(p.c.upper);

    // Uniform prior (simulation_impl.inl:523-537; estimate.jl:58-64, where a
    // fixed depth is one value shared by all components).
    std::uniform_real_distribution<double> um(p.m.lower, p.m.upper), uc(lowerC_, upperC_), uaz(p.az.lower, p.az.upper);
    m_.assign(C_, std::vector<double>(N_));
    log10c_.assign(C_, std::vector<double>(N_));
    az_.assign(Z_, std::vector<double>(N_));
    for (auto& row : m_)
        for (double& v : row) v = um(rng_);
    for (auto& row : log10c_)
        for (double& v : row) v = uc(rng_);
    for (auto& row : az_)
        for (double& v : row) v = uaz(rng_);
    dist_.assign(N_, 0.0);
    w_.assign(N_, 1.0 / (double)N_);
    compute_tau();
    gamma_ = p.adaptiveGamma ? -1.0 : p.gamma;

    // Model and observed histogram on the GPU.
    std::size_t maxF = *std::max_element(p.frameCounts.begin(), p.frameCounts.end());
    if (maxF < 1 || maxF > PCS_MAX_NK) throw std::invalid_argument("number_of_frames out of range");
    pcs_system sys{};
    sys.n_species = (uint32_t)C_;
    sys.kmin = (uint32_t)p.kmin;
    sys.n_k = (uint32_t)maxF;
    sys.n_de = (uint32_t)p.deBinCount;
    sys.de_bin_width = (float)(p.de.upper / (double)p.deBinCount);
    sys.four_dt = 4.0f * (float)p.deltaT;
    sys.half_l[0] = 0.5f * (float)p.Lx;
    sys.half_l[1] = 0.5f * (float)p.Ly;
    sys.half_l[2] = 0.5f * (float)p.Lz;
    sys.half_a[0] = 0.5f * (float)p.ax;
    sys.half_a[1] = 0.5f * (float)p.ay;
    sys.z_boundary = PCS_Z_PERIODIC;
    sys.detect_model = PCS_DETECT_SLAB;
    sys.dist_kind = PCS_DIST_L2SQ;
    std::vector<uint64_t> K(p.Ks.begin(), p.Ks.end());
    std::vector<uint32_t> H((std::size_t)sys.n_k * sys.n_de);
    uint64_t dropped = 0;
    check(pcs_reference_hist(K.data(), p.DEs.data(), K.size(), p.de.upper, sys.n_k, sys.n_de, H.data(), &dropped),
          "pcs_reference_hist");
    if (dropped && cfg.verbosity >= 0)
        std::fprintf(stderr, "Note: %llu observed tracks fall outside the histogram\n", (unsigned long long)dropped);
    check(pcs_open(&sys, H.data(), cfg.device, &ctx_), "pcs_open");
    const int r = pcs_batch_create(ctx_, &batch_);
    if (r != PCS_OK) {
        pcs_close(ctx_);
        check(r, "pcs_batch_create");
    }
}

Estimator::~Estimator()
{
    if (batch_) pcs_batch_destroy(batch_);
    if (ctx_) pcs_close(ctx_);
}

void Estimator::compute_tau()
{
    tauM_.resize(C_);
    tauC_.resize(C_);
    tauAz_.resize(Z_);
    for (std::size_t c = 0; c < C_; ++c) {
        tauM_[c] = std::sqrt(2.0 * var_1n(m_[c]));
        tauC_[c] = std::sqrt(2.0 * var_1n(log10c_[c]));
    }
    for (std::size_t z = 0; z < Z_; ++z) tauAz_[z] = std::sqrt(2.0 * var_1n(az_[z]));
}

Estimator::Trial Estimator::perturb(std::size_t parent)
{
    auto displace = [this](double x, double tau, double lo, double hi) {
        double d = 0.0;
        if (tau > 0.0) d = std::normal_distribution<double>(0.0, tau)(rng_);
        return reflect_into(x + d, lo, hi);
    };
    Trial t;
    t.m.resize(C_);
    t.log10c.resize(C_);
    t.az.resize(Z_);
    for (std::size_t c = 0; c < C_; ++c) {
        t.m[c] = displace(m_[c][parent], tauM_[c], par_.m.lower, par_.m.upper);
        t.log10c[c] = displace(log10c_[c][parent], tauC_[c], lowerC_, upperC_);
    }
    for (std::size_t z = 0; z < Z_; ++z) t.az[z] = displace(az_[z][parent], tauAz_[z], par_.az.lower, par_.az.upper);
    // Components ordered by m against label switching (simulation_impl.inl:274-297).
    std::vector<std::size_t> idx(C_);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&t](std::size_t a, std::size_t b) { return t.m[a] < t.m[b]; });
    Trial s = t;
    for (std::size_t c = 0; c < C_; ++c) {
        s.m[c] = t.m[idx[c]];
        s.log10c[c] = t.log10c[idx[c]];
        if (Z_ > 1) s.az[c] = t.az[idx[c]];
    }
    return s;
}

void Estimator::simulate(const std::vector<Trial>& trials, std::vector<double>& dist)
{
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    const std::size_t n = trials.size();
    const std::size_t J = par_.frameCounts.size();
    const double volume = par_.Lx * par_.Ly * par_.Lz / 1e12;
    std::vector<pcs_proposal> props(n);
    std::vector<uint32_t> frames(n * J), parts(n * J);
    for (std::size_t i = 0; i < n; ++i) {
        const Trial& t = trials[i];
        pcs_proposal& p = props[i];
        p = pcs_proposal{};
        p.sim_id = (uint32_t)i;
        p.first_video = (uint32_t)(i * J);
        p.n_videos = (uint32_t)J;
        // Per-trial precompute (simulation_impl.inl:306-341).
        std::vector<double> powc(C_);
        double csum = 0.0;
        for (std::size_t c = 0; c < C_; ++c) csum += powc[c] = std::pow(10.0, t.log10c[c]);
        double acc = 0.0;
        for (std::size_t c = 0; c < C_; ++c) {
            acc += powc[c] / csum;
            p.cum_prob[c] = (float)acc;
            p.sigma[c] = (float)std::sqrt(2.0 * t.m[c] * par_.deltaT);
            p.half_az[c] = 0.5f * (float)t.az[Z_ > 1 ? c : 0];
        }
        p.cum_prob[C_ - 1] = 2.0f;
        std::poisson_distribution<long long> pois(csum * volume);
        for (std::size_t v = 0; v < J; ++v) {
            frames[i * J + v] = (uint32_t)par_.frameCounts[v];
            const long long k = pois(rng_);
            parts[i * J + v] = (uint32_t)std::min<long long>(k, 0xFFFFFFFFll);
        }
    }
    dist.resize(n);
    autotune_once(props, frames, parts, J);
    const auto t1 = Clock::now();
    check(pcs_batch_load(batch_, props.data(), (uint32_t)n, frames.data(), parts.data(), (uint32_t)(n * J)),
          "pcs_batch_load");
    const auto t2 = Clock::now();
    check(pcs_batch_run(batch_, rng_(), cfg_.mode), "pcs_batch_run");
    check(pcs_batch_read(batch_, dist.data(), nullptr, nullptr), "pcs_batch_read");
    const auto t3 = Clock::now();
    if (cfg_.verbosity >= 2)
        std::printf("    round: %zu simulations, GPU run + read %.2f ms\n", n,
                    std::chrono::duration<double, std::milli>(t3 - t2).count());
    tPack_ += std::chrono::duration<double>(t1 - t0).count();
    tLoad_ += std::chrono::duration<double>(t2 - t1).count();
    tGpu_ += std::chrono::duration<double>(t3 - t2).count();
    simulations_ += n;
}

void Estimator::autotune_once(const std::vector<pcs_proposal>& props, const std::vector<uint32_t>& frames,
                              const std::vector<uint32_t>& parts, std::size_t J)
{
    if (tuned_ || !cfg_.autotune_min || props.size() < cfg_.autotune_min) return;
    tuned_ = true;
    // The first proposals hold videos [0, k J) in order. Enough of them for
    // about 5e9 particle-steps (some 10 ms a run: the tuner times 20 launch
    // shapes twice), at most 4000.
    std::size_t k = 0;
    double steps = 0.0;
    while (k < std::min<std::size_t>(props.size(), 4000) && steps < 5e9) {
        for (std::size_t v = 0; v < J; ++v) steps += (double)parts[k * J + v] * (double)frames[k * J + v];
        ++k;
    }
    pcs_tuning best{};
    float ms = 0.0f;
    check(pcs_autotune(ctx_, props.data(), (uint32_t)k, frames.data(), parts.data(), (uint32_t)(k * J), &best, &ms),
          "pcs_autotune");
    if (cfg_.verbosity >= 1)
        std::printf("  autotune: %u particle(s)/lane, %u waves/group, %u groups/CU (%.2f ms on %zu proposals)\n",
                    best.particles_per_lane, best.waves_per_group, best.groups_per_cu, ms, k);
}

void Estimator::simulate_proposed(std::vector<Trial>& trials, std::vector<double>& dist)
{
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    const std::size_t n = trials.size(), J = par_.frameCounts.size(), P = 2 * C_ + Z_;
    // population rows m_1..m_C, log10 c_1..log10 c_C, az_1..az_Z and their scales
    std::vector<double> theta(P * N_), cum(N_), tau(P);
    for (std::size_t c = 0; c < C_; ++c) {
        std::copy(m_[c].begin(), m_[c].end(), theta.begin() + (std::ptrdiff_t)(c * N_));
        std::copy(log10c_[c].begin(), log10c_[c].end(), theta.begin() + (std::ptrdiff_t)((C_ + c) * N_));
        tau[c] = tauM_[c];
        tau[C_ + c] = tauC_[c];
    }
    for (std::size_t z = 0; z < Z_; ++z) {
        std::copy(az_[z].begin(), az_[z].end(), theta.begin() + (std::ptrdiff_t)((2 * C_ + z) * N_));
        tau[2 * C_ + z] = tauAz_[z];
    }
    std::partial_sum(w_.begin(), w_.end(), cum.begin());
    pcs_prior pr{};
    pr.n_components = (uint32_t)C_;
    pr.n_depths = (uint32_t)Z_;
    pr.m_lo = par_.m.lower;
    pr.m_hi = par_.m.upper;
    pr.log10c_lo = lowerC_;
    pr.log10c_hi = upperC_;
    pr.az_lo = par_.az.lower;
    pr.az_hi = par_.az.upper;
    pr.deltat = par_.deltaT;
    pr.volume = par_.Lx * par_.Ly * par_.Lz / 1e12;
    std::vector<uint32_t> frames(n * J);
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t v = 0; v < J; ++v) frames[i * J + v] = (uint32_t)par_.frameCounts[v];
    std::vector<double> out(P * n);
    const auto t1 = Clock::now();
    check(pcs_batch_propose(batch_, &pr, theta.data(), cum.data(), tau.data(), (uint32_t)N_, (uint32_t)n,
                            frames.data(), (uint32_t)J, 0, rng_(), out.data()),
          "pcs_batch_propose");
    const auto t2 = Clock::now();
    for (std::size_t a = 0; a < n; ++a) {
        Trial& t = trials[a];
        t.m.resize(C_);
        t.log10c.resize(C_);
        t.az.resize(Z_);
        for (std::size_t c = 0; c < C_; ++c) {
            t.m[c] = out[c * n + a];
            t.log10c[c] = out[(C_ + c) * n + a];
        }
        for (std::size_t z = 0; z < Z_; ++z) t.az[z] = out[(2 * C_ + z) * n + a];
    }
    if (!tuned_ && cfg_.autotune_min && n >= cfg_.autotune_min) {
        // calibrate on the proposals just drawn (records and counts from the GPU)
        std::vector<pcs_proposal> props(n);
        std::vector<uint32_t> parts(n * J);
        check(pcs_batch_read_proposals(batch_, props.data(), parts.data()), "pcs_batch_read_proposals");
        autotune_once(props, frames, parts, J);
    }
    check(pcs_batch_run(batch_, rng_(), cfg_.mode), "pcs_batch_run");
    dist.resize(n);
    check(pcs_batch_read(batch_, dist.data(), nullptr, nullptr), "pcs_batch_read");
    const auto t3 = Clock::now();
    if (cfg_.verbosity >= 2)
        std::printf("    round: %zu proposals drawn on the GPU, propose + plan %.2f ms, run + read %.2f ms\n", n,
                    std::chrono::duration<double, std::milli>(t2 - t1).count(),
                    std::chrono::duration<double, std::milli>(t3 - t2).count());
    tPack_ += std::chrono::duration<double>(t1 - t0).count();
    tLoad_ += std::chrono::duration<double>(t2 - t1).count();
    tGpu_ += std::chrono::duration<double>(t3 - t2).count();
    simulations_ += n;
}

void Estimator::run()
{
    using Clock = std::chrono::steady_clock;
    converged_ = false;
    iterations_ = 0;
    if (gamma_ < 0.0) {
        if (cfg_.verbosity >= 0) std::printf("Compute initial gamma...\n");
        std::vector<Trial> prior(N_);
        for (std::size_t i = 0; i < N_; ++i) {
            prior[i].m.resize(C_);
            prior[i].log10c.resize(C_);
            prior[i].az.resize(Z_);
            for (std::size_t c = 0; c < C_; ++c) {
                prior[i].m[c] = m_[c][i];
                prior[i].log10c[c] = log10c_[c][i];
            }
            for (std::size_t z = 0; z < Z_; ++z) prior[i].az[z] = az_[z][i];
        }
        std::vector<double> d;
        simulate(prior, d);
        gamma_ = median_upper(d);
        if (cfg_.verbosity >= 0) std::printf("  initial gamma = %g\n", gamma_);
    }
    if (cfg_.verbosity >= 0) std::printf("Begin main iterations...\n");
    for (;;) {
        const auto t0 = Clock::now();
        ++iterations_;
        if (cfg_.maxIter && iterations_ > cfg_.maxIter) {
            std::printf("NOTE: --max-steps reached. stopping.\n");
            --iterations_;
            break;
        }
        gamma_ -= par_.deltaGamma;
        std::vector<double> cum(N_);
        std::partial_sum(w_.begin(), w_.end(), cum.begin());
        std::uniform_real_distribution<double> u01(0.0, 1.0);

        tPack_ = tLoad_ = tGpu_ = tWeights_ = 0.0;
        std::vector<Trial> cur(N_);
        std::vector<double> distBis(N_, std::numeric_limits<double>::infinity());
        std::vector<std::size_t> active(N_);
        std::iota(active.begin(), active.end(), 0);
        double trialsTotal = 0.0, trialsAvg = 0.0;
        std::size_t rounds = 0;
        while (!active.empty() && trialsAvg < (double)par_.avgTrialCount) {
            std::vector<Trial> batch(active.size());
            std::vector<double> d;
            if (cfg_.device_propose_min && active.size() >= cfg_.device_propose_min) {
                simulate_proposed(batch, d);
            } else {
                const auto tp = Clock::now();
                for (std::size_t a = 0; a < active.size(); ++a) {
                    const std::size_t parent = pick_by_weight(cum, u01(rng_));
                    batch[a] = perturb(parent);
                }
                tPack_ += std::chrono::duration<double>(Clock::now() - tp).count();
                simulate(batch, d);
            }
            std::vector<std::size_t> still;
            for (std::size_t a = 0; a < active.size(); ++a) {
                cur[active[a]] = std::move(batch[a]);
                distBis[active[a]] = d[a];
                if (d[a] > gamma_) still.push_back(active[a]);
            }
            trialsTotal += (double)active.size();
            trialsAvg = trialsTotal / (double)N_;
            active.swap(still);
            ++rounds;
        }
        const double secs = std::chrono::duration<double>(Clock::now() - t0).count();
        if (cfg_.verbosity >= 0)
            std::printf("Trials: %4.0f; gamma = %.4g, rounds %zu, %.2f s\n", trialsTotal, gamma_, rounds, secs);
        if (trialsAvg >= (double)par_.avgTrialCount) {
            // This generation never completed: keep the previous population
            // and report the tolerance it was accepted at (estimate.jl:237).
            converged_ = true;
            gamma_ += par_.deltaGamma;
            break;
        }
        // Weights from the new population against the previous one.
        const auto tw = Clock::now();
        if (par_.weightingScheme == io::WeightingScheme::pmcStandard) {
            // A parameter whose population has no spread (tau = 0: a prior
            // interval of one point) was never displaced; its kernel factor
            // is the same for every pair, so it is left out instead of
            // turning every weight into 0/0.
            std::vector<std::vector<double>> tn, to;
            std::vector<double> tau;
            auto add = [&](const std::vector<double>& row, const std::vector<double>& old, double t) {
                if (!(t > 0.0)) return;
                tn.push_back(row);
                to.push_back(old);
                tau.push_back(t);
            };
            for (std::size_t c = 0; c < C_; ++c) {
                std::vector<double> row(N_);
                for (std::size_t i = 0; i < N_; ++i) row[i] = cur[i].m[c];
                add(row, m_[c], tauM_[c]);
                for (std::size_t i = 0; i < N_; ++i) row[i] = cur[i].log10c[c];
                add(row, log10c_[c], tauC_[c]);
            }
            for (std::size_t z = 0; z < Z_; ++z) {
                std::vector<double> row(N_);
                for (std::size_t i = 0; i < N_; ++i) row[i] = cur[i].az[z];
                add(row, az_[z], tauAz_[z]);
            }
            if (tn.empty()) {
                w_.assign(N_, 1.0 / (double)N_);  // every member identical: equal weights
            } else if (N_ >= cfg_.device_weights_min) {
                // O(N^2) kernel sums on the GPU
                std::vector<double> xn, xo;
                for (std::size_t p = 0; p < tn.size(); ++p) {
                    xn.insert(xn.end(), tn[p].begin(), tn[p].end());
                    xo.insert(xo.end(), to[p].begin(), to[p].end());
                }
                std::vector<double> wn(N_);
                check(pcs_pmc_weights(cfg_.device, xn.data(), (uint32_t)N_, xo.data(), w_.data(), (uint32_t)N_,
                                      tau.data(), (uint32_t)tau.size(), wn.data()),
                      "pcs_pmc_weights");
                w_.swap(wn);
            } else {
                w_ = pmc_weights(tn, to, w_, tau);
            }
        } else {
            w_ = inverse_square_weights(distBis);
        }
        tWeights_ = std::chrono::duration<double>(Clock::now() - tw).count();
        if (cfg_.verbosity >= 1)
            std::printf("  host packing %.3f s, upload + plan %.3f s, GPU run + read %.3f s, weights %.3f s\n", tPack_,
                        tLoad_, tGpu_, tWeights_);
        for (std::size_t i = 0; i < N_; ++i) {
            for (std::size_t c = 0; c < C_; ++c) {
                m_[c][i] = cur[i].m[c];
                log10c_[c][i] = cur[i].log10c[c];
            }
            for (std::size_t z = 0; z < Z_; ++z) az_[z][i] = cur[i].az[z];
            dist_[i] = distBis[i];
        }
        compute_tau();
    }
}

io::Output Estimator::output() const
{
    io::Output o;
    o.model = par_.model;
    o.weightingScheme = par_.weightingScheme;
    o.abcSampleCount = N_;
    o.componentCount = C_;
    o.zComponentCount = Z_;
    o.converged = converged_;
    o.gamma = gamma_;
    for (std::size_t i = 0; i < N_; ++i) {
        for (std::size_t c = 0; c < C_; ++c) {
            o.m.push_back(m_[c][i]);
            o.c.push_back(std::pow(10.0, log10c_[c][i]));
        }
        for (std::size_t z = 0; z < Z_; ++z) o.az.push_back(az_[z][i]);
    }
    o.dist = dist_;
    o.w = w_;
    o.meta["simulator"] = "mi355x-hip";
    o.meta["iterations"] = std::to_string(iterations_);
    o.meta["simulations"] = std::to_string(simulations_);
    o.meta["gpus"] = "1";
    o.meta["gpu0_id"] = std::to_string(cfg_.device);
    o.meta["mode"] = cfg_.mode == PCS_MODE_PLAIN ? "plain" : "batched";
    o.meta["hostScalarType"] = "double";
    o.meta["deviceScalarType"] = "float";
    o.meta["distanceType"] = "exact-integer";
    o.number_of_iterations = (long long)iterations_;
    o.number_of_simulations = (long long)simulations_;
    return o;
}

}  // namespace pcs::abc
