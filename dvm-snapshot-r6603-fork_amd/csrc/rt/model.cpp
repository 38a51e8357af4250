// Learned cost model of the tuner.
//
// The reference has no such model: its tuner measures block shapes online
// (rts/dvmh/loop.cpp:1487-1909).  Here every timing the tuner records (loop
// site shape, candidate launch configuration, seconds) is a training sample
// for a small perceptron that predicts log(seconds per iteration point) from
// FP32 features of the site and the candidate.  Training runs on the GPU
// through the kernels of mlp.hip (forward, backward, Adam); ranking a site's
// candidates is one forward launch.  The tuner then measures only the model's
// best few candidates plus the default (tuner.cpp, mode 2); the model never
// touches loop arithmetic.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>

#include "tuner.h"
#include "../kernels/launch.h"

namespace dvmh {

namespace {

constexpr int F = DVMH_MLP_F, H = DVMH_MLP_H, NP = DVMH_MLP_NPARAM;
constexpr char MAGIC[8] = {'D', 'V', 'M', 'H', 'M', 'L', 'P', '1'};

struct Model {
    bool ready = false;
    std::vector<float> params;
    float mu[F], sd[F];
    float ymu = 0.f, ysd = 1.f;
    long samples = 0;
    double loss = 0;
};
Model model;
hipStream_t mstream = nullptr;

hipStream_t modelStream() {
    if (!mstream) HIP_CHECK(hipStreamCreateWithFlags(&mstream, hipStreamNonBlocking));
    return mstream;
}

double lg(double v) { return v > 0 ? std::log2(v) : 0.0; }

double sitePoints(const SiteInfo &s) {
    double p = 1;
    for (int i = 0; i < s.rank && i < 8; i++) p *= (double)std::max(1L, s.extents[i]);
    return p;
}

// Raw features of (site, candidate).
void features(const SiteInfo &s, const Candidate &c, float x[F]) {
    for (int f = 0; f < F; f++) x[f] = 0.f;
    if (s.kind >= 0 && s.kind <= 2) x[s.kind] = 1.f;
    if (s.kind == DVMH_AMD_KIND_STENCIL || s.kind == DVMH_AMD_KIND_BATCH || s.kind == DVMH_AMD_KIND_PAIR)
        x[DVMH_AMD_KIND_PARALLEL] = 1.f;   // a PARALLEL loop; the variant tells the kernel
    x[3] = (float)s.rank;
    x[4] = (float)s.depCount;
    for (int q = 0; q < 3; q++) {
        int ax = s.rank - 1 - q;   // innermost first
        x[7 - q] = ax >= 0 ? (float)lg((double)s.extents[ax]) : 0.f;
    }
    const double pts = sitePoints(s);
    x[8] = (float)lg(pts);
    const double threads = (double)c.block[0] * c.block[1] * c.block[2];
    const long depth = std::max(1L, (long)c.tune[DVMH_AMD_T_DEPTH]);
    x[9] = (float)lg(c.block[0]);
    x[10] = (float)lg(c.block[1]);
    x[11] = (float)lg(c.block[2]);
    x[12] = (float)lg((double)depth);
    x[13] = (float)c.tune[DVMH_AMD_T_XCD];
    x[14] = (float)c.tune[DVMH_AMD_T_VARIANT];
    x[15] = (float)lg(1.0 + (double)c.tune[DVMH_AMD_T_CHUNK]);
    x[16] = (float)lg(1.0 + (double)c.tune[DVMH_AMD_T_TILE_A]);
    x[17] = (float)lg(1.0 + (double)c.tune[DVMH_AMD_T_TILE_B]);
    x[18] = (float)c.tune[DVMH_AMD_T_TRANSFORM];
    x[19] = (float)lg(threads);
    // iteration points one workgroup covers per launch, and the resulting
    // number of workgroups (occupancy of 256 CUs)
    double perWg;
    if (c.tune[DVMH_AMD_T_VARIANT] == 1 && (s.kind == DVMH_AMD_KIND_STENCIL || s.kind == DVMH_AMD_KIND_PAIR)) {
        perWg = threads * (double)std::max(1L, (long)c.tune[DVMH_AMD_T_TILE_A]) * (double)std::max(1L, (long)c.tune[DVMH_AMD_T_CHUNK]);
    } else if (s.kind == DVMH_AMD_KIND_PARALLEL || s.kind == DVMH_AMD_KIND_STENCIL || s.kind == DVMH_AMD_KIND_BATCH) {
        perWg = threads * (double)depth;
    } else if (s.kind == DVMH_AMD_KIND_LINE) {
        double dep = 1;
        for (int i = 0; i < s.rank; i++) dep = std::max(dep, (double)s.extents[i]);
        perWg = threads * dep;   // each thread walks one line
    } else if (c.tune[DVMH_AMD_T_VARIANT] == 1) {
        // tiled wavefront: TILE_A x TILE_A x TILE_B iterations per workgroup
        perWg = (double)c.tune[DVMH_AMD_T_TILE_A] * (double)c.tune[DVMH_AMD_T_TILE_A] * (double)c.tune[DVMH_AMD_T_TILE_B];
    } else {
        // one hyperplane launch: a block of threads over the free dependent axes
        perWg = threads;
    }
    x[20] = (float)lg(perWg);
    x[21] = (float)lg(std::max(1.0, pts / perWg));
    // innermost extent against the block's x extent: partial waves on the row ends
    const double inner = s.rank > 0 ? (double)std::max(1L, s.extents[s.rank - 1]) : 1.0;
    x[22] = (float)(std::ceil(inner / c.block[0]) * c.block[0] / inner);
    x[23] = 1.f;
}

void standardize(float x[F]) {
    for (int f = 0; f < F; f++) x[f] = (x[f] - model.mu[f]) / model.sd[f];
}

// splitmix64 stream for the initial weights
uint64_t mix(uint64_t &st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float uniform(uint64_t &st, float a) { return (float)(((double)(mix(st) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * a); }

void initParams(std::vector<float> &p, uint64_t seed) {
    p.assign(NP, 0.f);
    uint64_t st = seed;
    const float a1 = std::sqrt(6.f / F), a2 = std::sqrt(6.f / H), a3 = std::sqrt(3.f / H);
    for (int i = 0; i < H * F; i++) p[i] = uniform(st, a1);
    for (int i = 0; i < H * H; i++) p[H * F + H + i] = uniform(st, a2);
    for (int i = 0; i < H; i++) p[H * F + H + H * H + H + i] = uniform(st, a3);
}

struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) { p = gpu.alloc(bytes ? bytes : 8); }
    ~DevBuf() { gpu.release(p); }
    float *f() const { return (float *)p; }
};

void up(void *d, const void *h, size_t bytes) {
    HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, modelStream()));
}
void down(void *h, const void *d, size_t bytes) {
    HIP_CHECK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, modelStream()));
    HIP_CHECK(hipStreamSynchronize(modelStream()));
}

void requireGpu(const char *what) {
    runtimeInit();
    DVMH_CHECK(gpu.present, "%s needs a GPU: the cost model trains and predicts with HIP kernels", what);
    DVMH_CHECK(!graphCapturing(), "%s cannot run during graph capture", what);
}

// Predicted seconds for every candidate of the site.
void predictSite(const SiteInfo &s, std::vector<double> &out) {
    const long n = (long)s.cands.size();
    std::vector<float> X((size_t)n * F);
    for (long i = 0; i < n; i++) {
        features(s, s.cands[i], &X[(size_t)i * F]);
        standardize(&X[(size_t)i * F]);
    }
    DevBuf dP(NP * sizeof(float)), dX(X.size() * sizeof(float)), dO((size_t)n * sizeof(float));
    up(dP.p, model.params.data(), NP * sizeof(float));
    up(dX.p, X.data(), X.size() * sizeof(float));
    dvmh_hip_mlp_forward(dP.f(), dX.f(), n, dO.f(), modelStream());
    std::vector<float> o((size_t)n);
    down(o.data(), dO.p, (size_t)n * sizeof(float));
    const double pts = sitePoints(s);
    out.resize((size_t)n);
    for (long i = 0; i < n; i++) out[i] = std::exp((double)o[i] * model.ysd + model.ymu) * pts;
}

const SiteInfo *findSite(const char *keyPart) {
    for (auto &kv : tunerSites())
        if (kv.first.find(keyPart) != std::string::npos) return &kv.second;
    return nullptr;
}

}  // namespace

// Ranking hook of the tuner: candidate indices by predicted time, fastest first.
int modelRank(SiteInfo &s, std::vector<int> &order) {
    order.clear();
    // chain sites are timed exhaustively: the records hold none of them
    if (!model.ready || !gpu.present || graphCapturing() || s.kind == DVMH_AMD_KIND_CHAIN) return 0;
    std::vector<double> pred;
    predictSite(s, pred);
    order.resize(pred.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pred[a] < pred[b]; });
    return (int)order.size();
}

// Trains a fresh network on every recorded timing (sites whose key contains
// `holdout` excluded).  Returns the number of samples.
long modelTrain(long epochs, double lr, uint64_t seed, const std::string &holdout, double *finalLoss) {
    requireGpu("Cost-model training");
    std::vector<float> X;
    std::vector<float> Y;
    for (auto &kv : tunerSites()) {
        const SiteInfo &s = kv.second;
        if (!holdout.empty() && kv.first.find(holdout) != std::string::npos) continue;
        const double pts = sitePoints(s);
        for (size_t c = 0; c < s.cands.size(); c++) {
            double t = s.measured((int)c);
            if (t <= 0) continue;
            float x[F];
            features(s, s.cands[c], x);
            X.insert(X.end(), x, x + F);
            Y.push_back((float)std::log(t / pts));
        }
    }
    const long n = (long)Y.size();
    if (n == 0) return 0;
    for (int f = 0; f < F; f++) {
        double m = 0, v = 0;
        for (long i = 0; i < n; i++) m += X[(size_t)i * F + f];
        m /= n;
        for (long i = 0; i < n; i++) v += (X[(size_t)i * F + f] - m) * (X[(size_t)i * F + f] - m);
        v /= n;
        model.mu[f] = (float)m;
        model.sd[f] = v > 1e-12 ? (float)std::sqrt(v) : 1.f;
    }
    double ym = 0, yv = 0;
    for (long i = 0; i < n; i++) ym += Y[i];
    ym /= n;
    for (long i = 0; i < n; i++) yv += (Y[i] - ym) * (Y[i] - ym);
    yv /= n;
    model.ymu = (float)ym;
    model.ysd = yv > 1e-12 ? (float)std::sqrt(yv) : 1.f;
    for (long i = 0; i < n; i++) {
        standardize(&X[(size_t)i * F]);
        Y[i] = (Y[i] - model.ymu) / model.ysd;
    }
    initParams(model.params, seed);
    DevBuf dP(NP * 4), dG(NP * 4), dM(NP * 4), dV(NP * 4), dX(X.size() * 4), dY((size_t)n * 4), dL(16),
            scratch((size_t)dvmh_hip_mlp_grad_scratch_bytes(n));
    std::vector<float> zeros(NP, 0.f);
    up(dP.p, model.params.data(), NP * 4);
    up(dM.p, zeros.data(), NP * 4);
    up(dV.p, zeros.data(), NP * 4);
    up(dX.p, X.data(), X.size() * 4);
    up(dY.p, Y.data(), (size_t)n * 4);
    const double b1 = (double)0.9f, b2 = (double)0.999f;   // FP32 betas, as the kernel uses them
    double p1 = 1, p2 = 1;
    for (long t = 1; t <= epochs; t++) {
        p1 *= b1;
        p2 *= b2;
        // cosine decay of the step size over the run
        double rate = lr * 0.5 * (1.0 + std::cos(M_PI * (double)(t - 1) / (double)epochs));
        dvmh_hip_mlp_grad(dP.f(), dX.f(), dY.f(), n, dG.f(), dL.f(), scratch.p, modelStream());
        dvmh_hip_mlp_adam(dP.f(), dG.f(), dM.f(), dV.f(), (float)rate, (float)b1, (float)b2, 1e-8f, (float)(1 - p1), (float)(1 - p2),
                modelStream());
    }
    // loss at the final parameters
    dvmh_hip_mlp_grad(dP.f(), dX.f(), dY.f(), n, dG.f(), dL.f(), scratch.p, modelStream());
    float loss = 0;
    down(model.params.data(), dP.p, NP * 4);
    down(&loss, dL.p, 4);
    model.ready = true;
    model.samples = n;
    model.loss = loss;
    if (finalLoss) *finalLoss = loss;
    for (auto &kv : tunerSites()) {
        kv.second.modelRanked = false;
        kv.second.modelOrder.clear();
    }
    return n;
}

long modelSave(const std::string &path) {
    if (!model.ready) return -1;
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return -1;
    int32_t dims[3] = {F, H, NP};
    bool ok = fwrite(MAGIC, 1, 8, f) == 8 && fwrite(dims, 4, 3, f) == 3 && fwrite(model.mu, 4, F, f) == (size_t)F &&
              fwrite(model.sd, 4, F, f) == (size_t)F && fwrite(&model.ymu, 4, 1, f) == 1 && fwrite(&model.ysd, 4, 1, f) == 1 &&
              fwrite(model.params.data(), 4, NP, f) == (size_t)NP;
    fclose(f);
    return ok ? NP : -1;
}

long modelLoad(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return -1;
    char magic[8];
    int32_t dims[3];
    Model m;
    m.params.resize(NP);
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, MAGIC, 8) == 0 && fread(dims, 4, 3, f) == 3 && dims[0] == F &&
              dims[1] == H && dims[2] == NP && fread(m.mu, 4, F, f) == (size_t)F && fread(m.sd, 4, F, f) == (size_t)F &&
              fread(&m.ymu, 4, 1, f) == 1 && fread(&m.ysd, 4, 1, f) == 1 && fread(m.params.data(), 4, NP, f) == (size_t)NP;
    fclose(f);
    if (!ok) return -1;
    m.ready = true;
    model = m;
    for (auto &kv : tunerSites()) {
        kv.second.modelRanked = false;
        kv.second.modelOrder.clear();
    }
    return NP;
}

void modelInit() {
    if (settings.modelFile.empty()) return;
    DVMH_CHECK(modelLoad(settings.modelFile) > 0, "Cannot load the cost model from %s (DVMH_AMD_MODEL)", settings.modelFile.c_str());
}

}  // namespace dvmh

using namespace dvmh;

extern "C" DvmType dvmh_amd_model_param_count_C(void) { return NP; }

extern "C" DvmType dvmh_amd_model_train_C(DvmType epochs, double lr, DvmType seed, const char *holdout, double *finalLoss) {
    runtimeInit();
    DVMH_CHECK(epochs >= 0 && lr > 0, "dvmh_amd_model_train_C: epochs must be >= 0 and lr > 0");
    return modelTrain(epochs, lr, (uint64_t)seed, holdout ? holdout : "", finalLoss);
}

extern "C" DvmType dvmh_amd_model_save_C(const char *path) { return modelSave(path); }
extern "C" DvmType dvmh_amd_model_load_C(const char *path) { return modelLoad(path); }
extern "C" DvmType dvmh_amd_model_ready_C(void) { return model.ready ? 1 : 0; }

extern "C" DvmType dvmh_amd_model_site_predict_C(const char *keyPart, double *pred, double *measured, DvmType maxCands) {
    runtimeInit();
    const SiteInfo *s = findSite(keyPart);
    if (!s) return -1;
    const long n = (long)s->cands.size();
    if (measured)
        for (long i = 0; i < n && i < maxCands; i++) measured[i] = s->measured((int)i);
    if (pred) {
        DVMH_CHECK(model.ready, "dvmh_amd_model_site_predict_C: no trained or loaded model");
        requireGpu("Cost-model prediction");
        std::vector<double> p;
        predictSite(*s, p);
        for (long i = 0; i < n && i < maxCands; i++) pred[i] = p[i];
    }
    return n;
}

extern "C" DvmType dvmh_amd_model_site_features_C(const char *keyPart, float *X, DvmType maxCands) {
    const SiteInfo *s = findSite(keyPart);
    if (!s) return -1;
    const long n = (long)s->cands.size();
    for (long i = 0; i < n && i < maxCands; i++) features(*s, s->cands[i], X + i * F);
    return n;
}

extern "C" DvmType dvmh_amd_model_grad_C(const float *params, const float *X, const float *y, DvmType n, float *grad, float *loss) {
    requireGpu("dvmh_amd_model_grad_C");
    DVMH_CHECK(n > 0, "dvmh_amd_model_grad_C: n must be positive");
    DevBuf dP(NP * 4), dG(NP * 4), dX((size_t)n * F * 4), dY((size_t)n * 4), dL(16), scratch((size_t)dvmh_hip_mlp_grad_scratch_bytes(n));
    up(dP.p, params, NP * 4);
    up(dX.p, X, (size_t)n * F * 4);
    up(dY.p, y, (size_t)n * 4);
    dvmh_hip_mlp_grad(dP.f(), dX.f(), dY.f(), n, dG.f(), dL.f(), scratch.p, modelStream());
    down(grad, dG.p, NP * 4);
    if (loss) down(loss, dL.p, 4);
    return NP;
}

extern "C" DvmType dvmh_amd_model_forward_C(const float *params, const float *X, DvmType n, float *out) {
    requireGpu("dvmh_amd_model_forward_C");
    DVMH_CHECK(n > 0, "dvmh_amd_model_forward_C: n must be positive");
    DevBuf dP(NP * 4), dX((size_t)n * F * 4), dO((size_t)n * 4);
    up(dP.p, params, NP * 4);
    up(dX.p, X, (size_t)n * F * 4);
    dvmh_hip_mlp_forward(dP.f(), dX.f(), n, dO.f(), modelStream());
    down(out, dO.p, (size_t)n * 4);
    return n;
}

extern "C" DvmType dvmh_amd_model_adam_C(float *params, const float *grad, float *m, float *v, DvmType step, double lr, double beta1,
        double beta2, double eps) {
    requireGpu("dvmh_amd_model_adam_C");
    DVMH_CHECK(step >= 1, "dvmh_amd_model_adam_C: step counts from 1");
    DevBuf dP(NP * 4), dG(NP * 4), dM(NP * 4), dV(NP * 4);
    up(dP.p, params, NP * 4);
    up(dG.p, grad, NP * 4);
    up(dM.p, m, NP * 4);
    up(dV.p, v, NP * 4);
    // the kernel computes with FP32 betas; the bias corrections use the same values
    const double fb1 = (double)(float)beta1, fb2 = (double)(float)beta2;
    dvmh_hip_mlp_adam(dP.f(), dG.f(), dM.f(), dV.f(), (float)lr, (float)fb1, (float)fb2, (float)eps,
            (float)(1.0 - std::pow(fb1, (double)step)), (float)(1.0 - std::pow(fb2, (double)step)), modelStream());
    down(params, dP.p, NP * 4);
    down(m, dM.p, NP * 4);
    down(v, dV.p, NP * 4);
    return NP;
}

extern "C" DvmType dvmh_amd_tuner_sites_C(char *buf, DvmType bufLen) {
    std::string all;
    for (auto &kv : tunerSites()) {
        all += kv.first;
        all += '\n';
    }
    if (buf && bufLen > 0) {
        size_t k = std::min((size_t)(bufLen - 1), all.size());
        memcpy(buf, all.data(), k);
        buf[k] = 0;
    }
    return (DvmType)all.size() + 1;
}
