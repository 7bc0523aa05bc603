#include "estimator.h"

#include "../execution/hip_generic_executor.h"
#include "../execution/runtime_compiler.h"
#include "../infra/generators.h"
#include "../planning/hip_plan_builder.h"
#include "gdb/gdb_kernels.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <cstring>
#include <fstream>
#include <numeric>
#include <stdexcept>
#include <climits>
#include <vector>

namespace gdb {
namespace runtime {

namespace {

void check(gdb_status st, const char* what) {
    if (st != GDB_OK) {
        const char* m = gdb_last_error();
        if (!m || !*m) m = gdb_kernels_last_error();
        throw std::runtime_error(std::string(what) + ": " + (m ? m : ""));
    }
}

double lg2(double x) { return std::log2(std::max(1.0, x)); }
float clampf(double x, double lo, double hi) { return static_cast<float>(std::min(hi, std::max(lo, x))); }

struct DevBuf {
    gdb_ctx ctx;
    void* p = nullptr;
    DevBuf(gdb_ctx c, size_t bytes) : ctx(c) { check(gdb_alloc(c, bytes, &p), "gdb_alloc"); }
    ~DevBuf() { gdb_free(ctx, p); }
    template <typename T>
    T* as() {
        return static_cast<T*>(p);
    }
};

}  // namespace

double targetOf(double groups, double ub) { return lg2(groups) - lg2(ub); }
double groupsFromTarget(double t, double ub) { return std::max(1.0, std::exp2(t) * std::max(1.0, ub)); }

FeatureOut computeFeatures(const FeatureInput& in) {
    FeatureOut out;
    const double n = static_cast<double>(in.nRows);
    const double range = static_cast<double>(std::max<int64_t>(1, in.keyMax - in.keyMin + 1));
    // Key frequencies of the (filtered) sample in an open-addressing table
    // at most half full, then the frequencies' own histogram: the features
    // need only counts of counts (singletons, pairs, the 8 largest, entropy),
    // so each distinct count costs one log2 instead of each distinct key.
    thread_local std::vector<int64_t> tkey;
    thread_local std::vector<uint32_t> tcnt, mult;
    size_t cap = 16;
    while (cap < 2 * static_cast<size_t>(in.sampleSize) + 2) cap <<= 1;
    tkey.assign(cap, INT64_MIN);
    tcnt.assign(cap, 0u);
    uint64_t s = 0;
    size_t d = 0;
    for (uint64_t i = 0; i < in.sampleSize; ++i) {
        if (in.sampleVals && in.sampleVals[i] >= in.valCut) continue;
        const int64_t k = in.sampleKeys[i];
        size_t h = (static_cast<uint64_t>(static_cast<uint32_t>(k)) * 0x9E3779B97F4A7C15ull) >> 20;
        for (h &= cap - 1; tkey[h] != INT64_MIN && tkey[h] != k; h = (h + 1) & (cap - 1)) {
        }
        if (tkey[h] == INT64_MIN) {
            tkey[h] = k;
            ++d;
        }
        ++tcnt[h];
        ++s;
    }
    const double S = static_cast<double>(std::max<uint64_t>(1, in.sampleSize));
    const double sel = in.sampleSize ? static_cast<double>(s) / S : 1.0;
    const double nEst = std::max(1.0, n * sel);
    const double ub = std::max(1.0, std::min(nEst, range));
    mult.assign(static_cast<size_t>(s) + 1, 0u);  // mult[c] = keys seen c times
    uint32_t maxc = 0;
    for (size_t h = 0; h < cap; ++h)
        if (tcnt[h]) {
            ++mult[tcnt[h]];
            maxc = std::max(maxc, tcnt[h]);
        }
    const double f1 = s >= 1 ? mult[1] : 0, f2 = s >= 2 ? mult[2] : 0;
    const double ss = std::max(1.0, static_cast<double>(s));
    double top1 = maxc, top8 = 0, ent = 0;
    uint32_t taken = 0;
    for (uint32_t c = maxc; c >= 1; --c) {
        if (!mult[c]) continue;
        const uint32_t m = std::min<uint32_t>(mult[c], 8u - std::min(8u, taken));
        top8 += static_cast<double>(m) * c;
        taken += m;
        const double p = c / ss;
        ent -= static_cast<double>(mult[c]) * p * std::log2(p);
    }
    const double dd = static_cast<double>(d);
    const double gee = std::sqrt(nEst / ss) * f1 + (dd - f1);
    const double chao = dd + (f2 > 0 ? f1 * f1 / (2.0 * f2) : f1 * (f1 - 1.0) / 2.0);
    Features& x = out.x;
    x[0] = static_cast<float>(lg2(n) / 32.0);
    x[1] = static_cast<float>(lg2(range) / 32.0);
    x[2] = static_cast<float>(sel);
    x[3] = static_cast<float>(lg2(nEst) / 32.0);
    x[4] = static_cast<float>(dd / ss);
    x[5] = static_cast<float>(f1 / ss);
    x[6] = static_cast<float>(f2 / ss);
    x[7] = static_cast<float>(top1 / ss);
    x[8] = static_cast<float>(top8 / ss);
    x[9] = static_cast<float>(lg2(gee) / 32.0);
    x[10] = static_cast<float>(lg2(chao) / 32.0);
    x[11] = static_cast<float>(lg2(ub) / 32.0);
    x[12] = static_cast<float>(lg2(dd) / 32.0);
    x[13] = static_cast<float>(ss > 1 ? ent / std::log2(ss) : 0.0);
    x[14] = clampf((lg2(nEst) - lg2(range)) / 10.0, -1.0, 1.0);
    x[15] = clampf((lg2(std::min(gee, ub)) - lg2(ub)) / 10.0, -2.0, 0.0);
    out.upperBound = ub;
    out.sampleSel = sel;
    return out;
}

void sampleDeviceColumns(gdb_ctx ctx, gdb_stream s, const int32_t* dkeys, const int64_t* dvals, uint64_t n,
                         uint64_t sampleSize, int32_t* hkeys, int64_t* hvals) {
    (void)ctx;
    if (sampleSize > n) sampleSize = n;
    static std::mutex mu;
    static uint8_t* stage = nullptr;  // pinned, kept until the process ends
    static size_t stageBytes = 0;
    std::lock_guard<std::mutex> lock(mu);
    const size_t need = static_cast<size_t>(sampleSize) * 12 + 16;
    if (need > stageBytes) {
        check(gdb_stream_sync(s), "sample sync");
        if (stage) gdb_host_free_pinned(stage);
        stage = nullptr;
        stageBytes = 0;
        void* h = nullptr;
        check(gdb_host_alloc_pinned(std::max<size_t>(need, 1 << 16), &h), "gdb_host_alloc_pinned");
        stage = static_cast<uint8_t*>(h);
        stageBytes = std::max<size_t>(need, 1 << 16);
    }
    int32_t* sk = reinterpret_cast<int32_t*>(stage);
    int64_t* sv = reinterpret_cast<int64_t*>(stage + ((static_cast<size_t>(sampleSize) * 4 + 15) & ~static_cast<size_t>(15)));
    const uint64_t chunks = sampleSize >= 8 ? 8 : 1;
    const uint64_t per = sampleSize / chunks;
    uint64_t done = 0;
    for (uint64_t c = 0; c < chunks; ++c) {
        const uint64_t len = (c + 1 == chunks) ? sampleSize - done : per;
        const uint64_t start = (n - len) * c / std::max<uint64_t>(1, chunks - 1);
        check(gdb_memcpy_d2h_async(sk + done, dkeys + start, len * 4, s), "sample keys");
        if (dvals && hvals) check(gdb_memcpy_d2h_async(sv + done, dvals + start, len * 8, s), "sample vals");
        done += len;
    }
    check(gdb_stream_sync(s), "sample sync");
    std::memcpy(hkeys, sk, static_cast<size_t>(sampleSize) * 4);
    if (dvals && hvals) std::memcpy(hvals, sv, static_cast<size_t>(sampleSize) * 8);
}

// ---------------------------------------------------------------------------

CardinalityEstimator::CardinalityEstimator(gdb_ctx ctx) : ctx_(ctx), stream_(gdb_default_stream(ctx)) {
    params_.assign(GDB_MLP_NPARAMS, 0.0f);
    mean_.fill(0.0f);
    std_.fill(1.0f);
    initWeights(1);
}

CardinalityEstimator::~CardinalityEstimator() {
    if (dP_) gdb_free(ctx_, dP_);
    if (dX_) gdb_free(ctx_, dX_);
    if (dY_) gdb_free(ctx_, dY_);
    if (hStage_) gdb_host_free_pinned(hStage_);
}

void CardinalityEstimator::reserve(int n) {
    if (n <= cap_ && dP_) return;
    check(gdb_stream_sync(stream_), "sync");
    const int cap = std::max(n, 64);
    for (void** p : {&dP_, &dX_, &dY_})
        if (*p) gdb_free(ctx_, *p);
    if (hStage_) gdb_host_free_pinned(hStage_);
    dP_ = dX_ = dY_ = nullptr;
    hStage_ = nullptr;
    cap_ = 0;
    check(gdb_alloc(ctx_, GDB_MLP_NPARAMS * 4, &dP_), "gdb_alloc");
    check(gdb_alloc(ctx_, static_cast<size_t>(cap) * kFeatures * 4, &dX_), "gdb_alloc");
    check(gdb_alloc(ctx_, static_cast<size_t>(cap) * 4, &dY_), "gdb_alloc");
    void* h = nullptr;
    check(gdb_host_alloc_pinned((GDB_MLP_NPARAMS + static_cast<size_t>(cap) * (kFeatures + 1)) * 4, &h),
          "gdb_host_alloc_pinned");
    hStage_ = static_cast<float*>(h);
    cap_ = cap;
}

void CardinalityEstimator::initWeights(uint64_t seed) {
    auto normal = [&](uint64_t i) {
        const double u1 = std::max(1e-12, infra::uniform01(infra::rng(seed, 901, 2 * i)));
        const double u2 = infra::uniform01(infra::rng(seed, 901, 2 * i + 1));
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    std::fill(params_.begin(), params_.end(), 0.0f);
    for (int i = 0; i < 1024; ++i) params_[static_cast<size_t>(i)] = static_cast<float>(normal(i) * std::sqrt(2.0 / 16.0));
    for (int i = 0; i < 4096; ++i)
        params_[static_cast<size_t>(1088 + i)] = static_cast<float>(normal(10000 + i) * std::sqrt(2.0 / 64.0));
    for (int i = 0; i < 64; ++i)
        params_[static_cast<size_t>(5248 + i)] = static_cast<float>(normal(20000 + i) * std::sqrt(1.0 / 64.0));
}

void CardinalityEstimator::standardise(const float* X, int n, std::vector<float>& out) const {
    out.resize(static_cast<size_t>(n) * kFeatures);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < kFeatures; ++k)
            out[static_cast<size_t>(i) * kFeatures + k] =
                (X[static_cast<size_t>(i) * kFeatures + k] - mean_[k]) / std_[k];
}

void CardinalityEstimator::train(const float* X, const float* y, int n, const TrainConfig& cfg,
                                 std::vector<float>* losses) {
    if (n <= 0) throw std::runtime_error("estimator: no training samples");
    for (int k = 0; k < kFeatures; ++k) {
        double m = 0, v = 0;
        for (int i = 0; i < n; ++i) m += X[static_cast<size_t>(i) * kFeatures + k];
        m /= n;
        for (int i = 0; i < n; ++i) {
            const double d = X[static_cast<size_t>(i) * kFeatures + k] - m;
            v += d * d;
        }
        mean_[k] = static_cast<float>(m);
        std_[k] = static_cast<float>(std::sqrt(v / n) > 1e-6 ? std::sqrt(v / n) : 1.0);
    }
    std::vector<float> Xs;
    standardise(X, n, Xs);
    const int B = std::max(1, std::min(cfg.batch, n));
    DevBuf dP(ctx_, GDB_MLP_NPARAMS * 4), dM(ctx_, GDB_MLP_NPARAMS * 4), dV(ctx_, GDB_MLP_NPARAMS * 4);
    DevBuf dG(ctx_, static_cast<size_t>(B) * GDB_MLP_NPARAMS * 4), dL(ctx_, static_cast<size_t>(B) * 4);
    DevBuf dXb(ctx_, static_cast<size_t>(B) * kFeatures * 4), dYb(ctx_, static_cast<size_t>(B) * 4);
    DevBuf dXall(ctx_, static_cast<size_t>(n) * kFeatures * 4), dYall(ctx_, static_cast<size_t>(n) * 4);
    check(gdb_memcpy_h2d_async(dP.p, params_.data(), GDB_MLP_NPARAMS * 4, stream_), "upload params");
    const uint8_t z = 0;
    check(gdb_fill_async(dM.p, GDB_MLP_NPARAMS * 4, &z, 1, stream_), "zero m");
    check(gdb_fill_async(dV.p, GDB_MLP_NPARAMS * 4, &z, 1, stream_), "zero v");
    check(gdb_memcpy_h2d_async(dXall.p, Xs.data(), Xs.size() * 4, stream_), "upload X");
    std::vector<int> perm(static_cast<size_t>(n));
    std::vector<float> xb(static_cast<size_t>(B) * kFeatures), yb(static_cast<size_t>(B)), yhat(static_cast<size_t>(n));
    int step = 0;
    for (int ep = 0; ep < cfg.epochs; ++ep) {
        std::iota(perm.begin(), perm.end(), 0);
        for (int i = n - 1; i > 0; --i) {
            const int j = static_cast<int>(infra::uniformInt(infra::rng(cfg.seed, 77 + static_cast<uint64_t>(ep), static_cast<uint64_t>(i)), 0, i));
            std::swap(perm[static_cast<size_t>(i)], perm[static_cast<size_t>(j)]);
        }
        for (int b0 = 0; b0 < n; b0 += B) {
            const int nb = std::min(B, n - b0);
            for (int r = 0; r < nb; ++r) {
                const int src = perm[static_cast<size_t>(b0 + r)];
                std::memcpy(&xb[static_cast<size_t>(r) * kFeatures], &Xs[static_cast<size_t>(src) * kFeatures], kFeatures * 4);
                yb[static_cast<size_t>(r)] = y[src];
            }
            check(gdb_memcpy_h2d_async(dXb.p, xb.data(), static_cast<size_t>(nb) * kFeatures * 4, stream_), "batch X");
            check(gdb_memcpy_h2d_async(dYb.p, yb.data(), static_cast<size_t>(nb) * 4, stream_), "batch y");
            check(gdb_mlp_backward(dP.as<float>(), dXb.as<float>(), dYb.as<float>(), dG.as<float>(), dL.as<float>(), nb,
                                   1.0f / static_cast<float>(nb), stream_),
                  "mlp backward");
            check(gdb_mlp_reduce_adam(dP.as<float>(), dM.as<float>(), dV.as<float>(), dG.as<float>(), nullptr, nb, ++step,
                                      cfg.lr, cfg.beta1, cfg.beta2, cfg.eps, stream_),
                  "mlp adam");
        }
        if (losses) {
            check(gdb_mlp_forward(dP.as<float>(), dXall.as<float>(), dYall.as<float>(), n, stream_), "mlp forward");
            check(gdb_memcpy_d2h_async(yhat.data(), dYall.p, static_cast<size_t>(n) * 4, stream_), "download");
            check(gdb_stream_sync(stream_), "sync");
            double mse = 0;
            for (int i = 0; i < n; ++i) mse += (yhat[static_cast<size_t>(i)] - y[i]) * (yhat[static_cast<size_t>(i)] - y[i]);
            losses->push_back(static_cast<float>(mse / n));
        }
    }
    check(gdb_memcpy_d2h_async(params_.data(), dP.p, GDB_MLP_NPARAMS * 4, stream_), "download params");
    check(gdb_stream_sync(stream_), "sync");
    loaded_ = true;
}

void CardinalityEstimator::predict(const float* X, int n, float* out) {
    if (n <= 0) return;
    std::vector<float> Xs;
    standardise(X, n, Xs);
    reserve(n);
    // weights, inputs and outputs staged in one pinned block: DMA copies,
    // no allocation, one synchronisation
    float* hP = hStage_;
    float* hX = hStage_ + GDB_MLP_NPARAMS;
    float* hY = hX + static_cast<size_t>(cap_) * kFeatures;
    std::memcpy(hP, params_.data(), GDB_MLP_NPARAMS * 4);
    std::memcpy(hX, Xs.data(), Xs.size() * 4);
    check(gdb_memcpy_h2d_async(dP_, hP, GDB_MLP_NPARAMS * 4, stream_), "upload params");
    check(gdb_memcpy_h2d_async(dX_, hX, Xs.size() * 4, stream_), "upload X");
    check(gdb_mlp_forward(static_cast<const float*>(dP_), static_cast<const float*>(dX_), static_cast<float*>(dY_), n,
                          stream_),
          "mlp forward");
    check(gdb_memcpy_d2h_async(hY, dY_, static_cast<size_t>(n) * 4, stream_), "download");
    check(gdb_stream_sync(stream_), "sync");
    std::memcpy(out, hY, static_cast<size_t>(n) * 4);
}

double CardinalityEstimator::predictGroups(const FeatureOut& f) {
    float t = 0.0f;
    predict(f.x.data(), 1, &t);
    return groupsFromTarget(t, f.upperBound);
}

namespace {
constexpr char kMagic[8] = {'G', 'D', 'B', 'M', 'L', 'P', '0', '1'};
}

void CardinalityEstimator::save(const std::string& path) const {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) throw std::runtime_error("estimator: cannot write " + path);
    const uint32_t np = GDB_MLP_NPARAMS, nf = kFeatures;
    f.write(kMagic, 8);
    f.write(reinterpret_cast<const char*>(&np), 4);
    f.write(reinterpret_cast<const char*>(&nf), 4);
    f.write(reinterpret_cast<const char*>(mean_.data()), kFeatures * 4);
    f.write(reinterpret_cast<const char*>(std_.data()), kFeatures * 4);
    f.write(reinterpret_cast<const char*>(params_.data()), GDB_MLP_NPARAMS * 4);
}

void CardinalityEstimator::load(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("estimator: cannot read " + path);
    char magic[8];
    uint32_t np = 0, nf = 0;
    f.read(magic, 8);
    f.read(reinterpret_cast<char*>(&np), 4);
    f.read(reinterpret_cast<char*>(&nf), 4);
    if (!f || std::memcmp(magic, kMagic, 8) != 0 || np != GDB_MLP_NPARAMS || nf != kFeatures)
        throw std::runtime_error("estimator: " + path + " is not a weights file of this network");
    f.read(reinterpret_cast<char*>(mean_.data()), kFeatures * 4);
    f.read(reinterpret_cast<char*>(std_.data()), kFeatures * 4);
    f.read(reinterpret_cast<char*>(params_.data()), GDB_MLP_NPARAMS * 4);
    if (!f) throw std::runtime_error("estimator: truncated weights file " + path);
    loaded_ = true;
}

// ---------------------------------------------------------------------------
// Training runs labelled by the engine's exact distinct-count plan.
// ---------------------------------------------------------------------------
std::vector<RunLog> collectRuns(gdb_ctx ctx, int nruns, uint64_t seed, uint64_t nmin, uint64_t nmax, int64_t kmin,
                                int64_t kmax, double smax, double selmin) {
    using namespace codegen;
    auto plan = buildHipPlan("GB_COUNT", parsePlanOptions("filter=1"));
    HipCodegen cg = generateFromPlan(*plan);
    RuntimeCompiler rc(ctx);
    auto cq = rc.compile(cg.print());
    HipGenericExecutor ex(ctx);
    for (const auto& bt : plan->bufferTables) ex.registerBufferTable(bt.table, bt.column, bt.buffer, bt.rowsExpr);
    DevBuf keys(ctx, nmax * 4 + 256), vals(ctx, nmax * 8 + 256);
    gdb_stream s = ex.stream();
    constexpr uint64_t kSample = 8192;
    std::vector<int32_t> hk(kSample);
    std::vector<int64_t> hv(kSample);
    std::vector<RunLog> logs;
    logs.reserve(static_cast<size_t>(nruns));
    for (int r = 0; r < nruns; ++r) {
        auto u = [&](uint64_t k) { return infra::uniform01(infra::rng(seed, 500 + k, static_cast<uint64_t>(r))); };
        RunLog log;
        log.spec.n = static_cast<uint64_t>(std::llround(std::exp2(std::log2(static_cast<double>(nmin)) +
                                                                   u(0) * (std::log2(static_cast<double>(nmax)) -
                                                                           std::log2(static_cast<double>(nmin))))));
        log.spec.n = std::min(nmax, std::max(nmin, log.spec.n));
        log.spec.nkeys = static_cast<int64_t>(std::llround(std::exp2(std::log2(static_cast<double>(kmin)) +
                                                                      u(1) * (std::log2(static_cast<double>(kmax)) -
                                                                              std::log2(static_cast<double>(kmin))))));
        log.spec.skew = u(2) * smax;
        log.spec.sel = selmin + u(3) * (1.0 - selmin);
        log.spec.seed = infra::rng(seed, 600, static_cast<uint64_t>(r));
        const int64_t cut = static_cast<int64_t>(std::floor(log.spec.sel * 1000000.0)) + 1;
        check(gdb_gen_device_zipf_i32(keys.as<int32_t>(), log.spec.n, log.spec.skew, log.spec.nkeys, log.spec.seed, s),
              "device zipf");
        check(gdb_gen_device_uniform_i64(vals.as<int64_t>(), log.spec.n, 0, 1000000, log.spec.seed, s), "device vals");
        const uint64_t smp = std::min<uint64_t>(kSample, log.spec.n);
        sampleDeviceColumns(ctx, s, keys.as<int32_t>(), vals.as<int64_t>(), log.spec.n, smp, hk.data(), hv.data());
        FeatureInput fi;
        fi.nRows = log.spec.n;
        fi.keyMin = 1;
        fi.keyMax = log.spec.nkeys;
        fi.valCut = cut;
        fi.sampleKeys = hk.data();
        fi.sampleVals = hv.data();
        fi.sampleSize = smp;
        log.f = computeFeatures(fi);
        ex.registerTable("G", log.spec.n);
        ex.registerTableColumn("G", "g_key", keys.p);
        ex.registerTableColumn("G", "g_val", vals.p);
        ex.registerScalarInt("gb_key_min", 1);
        ex.registerScalarInt("gb_val_cut", cut);
        ex.registerSymbol("gb_key_bits", static_cast<uint64_t>(log.spec.nkeys));
        ex.execute(rc, cq, cg, 0, 1);
        uint64_t distinct = 0;
        ex.readBuffer("d_gb_distinct", 0, &distinct, 8);
        log.groups = static_cast<double>(distinct);
        logs.push_back(log);
    }
    return logs;
}

}  // namespace runtime
}  // namespace gdb
