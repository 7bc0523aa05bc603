#pragma once
// Group-count (cardinality) estimator for hash group-by sizing.
//
// No counterpart in the reference, which sizes every table by an exact
// max-key pre-pass (codegen/execution/max_key_symbols.cpp:188-235). Here an
// FP32 MLP (16 features -> 64 -> 64 -> 1, ReLU) predicts the number of groups
// from plan/column statistics and a small row sample; the runner sizes the
// hash table from the prediction and picks the LDS or global variant. The
// network is trained on the GPU by the fixed kernels of gdb_kernels.h on
// (features, true count) pairs that the engine logs from its own exact-count
// runs. The exact pre-pass plus resize-on-overflow stays the fallback and the
// judge of the estimate.

#include "gdb/gdb_hip.h"

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace gdb {
namespace runtime {

constexpr int kFeatures = 16;
using Features = std::array<float, kFeatures>;

// Statistics the planner has before running a group-by: catalog ranges plus
// a row sample (keys and values, already on the host).
struct FeatureInput {
    uint64_t nRows = 0;
    int64_t keyMin = 0, keyMax = 0;
    int64_t valCut = INT64_MAX;  // rows qualify when val < valCut
    const int32_t* sampleKeys = nullptr;
    const int64_t* sampleVals = nullptr;
    uint64_t sampleSize = 0;
};

struct FeatureOut {
    Features x{};
    double upperBound = 1.0;   // min(estimated qualifying rows, key range)
    double sampleSel = 1.0;
};

FeatureOut computeFeatures(const FeatureInput& in);

// Target transform: t = log2(groups) - log2(upperBound); groups = ub * 2^t.
double targetOf(double groups, double ub);
double groupsFromTarget(double t, double ub);

// One logged training run: the generator's parameters, the features the
// planner would compute, and the true group count from the engine's exact
// distinct-count plan.
struct RunSpec {
    uint64_t n = 0;
    double skew = 0.0;
    int64_t nkeys = 0;
    double sel = 1.0;
    uint64_t seed = 0;
};
struct RunLog {
    RunSpec spec;
    FeatureOut f;
    double groups = 0.0;
};
// Draw `nruns` generator settings (n log-uniform in [nmin, nmax], key range
// log-uniform in [kmin, kmax], skew uniform in [0, smax], selectivity uniform
// in [selmin, 1]), generate each on the device, compute features from a
// sample and label it with the exact count.
std::vector<RunLog> collectRuns(gdb_ctx ctx, int nruns, uint64_t seed, uint64_t nmin, uint64_t nmax, int64_t kmin,
                                int64_t kmax, double smax, double selmin);

struct TrainConfig {
    int epochs = 300;
    int batch = 256;
    float lr = 1e-3f;
    float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    uint64_t seed = 1;
};

class CardinalityEstimator {
public:
    explicit CardinalityEstimator(gdb_ctx ctx);
    ~CardinalityEstimator();
    CardinalityEstimator(const CardinalityEstimator&) = delete;
    CardinalityEstimator& operator=(const CardinalityEstimator&) = delete;

    // He-initialised weights from a seed (host side, deterministic).
    void initWeights(uint64_t seed);
    // Train on the GPU; fits the feature standardisation first. losses (one
    // mean loss per epoch) is optional.
    void train(const float* X, const float* y, int n, const TrainConfig& cfg, std::vector<float>* losses);
    // GPU inference of the (standardised-input) network: out[i] = target.
    void predict(const float* X, int n, float* out);
    double predictGroups(const FeatureOut& f);

    void save(const std::string& path) const;
    void load(const std::string& path);
    bool loaded() const { return loaded_; }

    std::vector<float>& params() { return params_; }
    std::array<float, kFeatures>& mean() { return mean_; }
    std::array<float, kFeatures>& stdev() { return std_; }

private:
    gdb_ctx ctx_;
    gdb_stream stream_;
    std::vector<float> params_;
    std::array<float, kFeatures> mean_{}, std_{};
    bool loaded_ = false;
    // inference buffers kept across calls (no allocation per prediction):
    // device weights, inputs and outputs for up to cap_ samples, and one
    // pinned host block staging all three copies
    void* dP_ = nullptr;
    void* dX_ = nullptr;
    void* dY_ = nullptr;
    float* hStage_ = nullptr;
    int cap_ = 0;
    void reserve(int n);
    void standardise(const float* X, int n, std::vector<float>& out) const;
};

// Copy a strided sample of `sampleSize` rows (8 evenly spaced contiguous
// chunks) of a device key column (and optional value column) to the host,
// through a process-wide pinned staging block (grown on demand and kept
// until the process ends) so the copies are DMA and one synchronisation.
void sampleDeviceColumns(gdb_ctx ctx, gdb_stream s, const int32_t* dkeys, const int64_t* dvals, uint64_t n,
                         uint64_t sampleSize, int32_t* hkeys, int64_t* hvals);

}  // namespace runtime
}  // namespace gdb
