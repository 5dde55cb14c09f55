// spg_predict.hip - learned per-row output-nnz predictor: a 4-64-64-1 FP32
// MLP over cheap row features, with hand-written inference and training
// kernels (no reference counterpart; the reference sizes nothing).
//
// Features of C row i (written by InitGEMM): log1p of nnz(A row i), of the
// flop upper bound ub_i = sum over k of nnz(B row k), of max_k nnz(B row k)
// and of the standard deviation of those nnz(B row k).  Target: the exact
// count from the symbolic pass, as t = log1p(nnz(C row i)) - log1p(s_i),
// where s_i = n * (1 - exp(-ub_i / n)) is the count expected if the row's
// ub_i products fell uniformly on B's n columns; the network learns the
// correction for skew and repeated columns, which keeps it usable on
// products with other n than it was trained on.  The prediction is
// expm1(y + log1p(s_i)) clamped to [0, min(ub_i, n)].
//
// Training is deterministic: a workgroup computes the per-sample
// activations and deltas of 64 samples into LDS, then every thread sums the
// gradient of its own parameters over those samples in a fixed order; the
// per-workgroup partials are reduced in workgroup order by the Adam kernel.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "spg_internal.h"
#include "spg_rng.h"

namespace spg {

constexpr int kF = 4;    // features
constexpr int kH = 64;   // hidden width
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kF;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kNP = kOffB3 + 1;  // 4545 parameters
constexpr int kSamplesPerWG = 64;

__device__ __forceinline__ float relu(float x) { return x > 0.f ? x : 0.f; }

// log1p of the distinct count expected for u products uniform on n columns.
__device__ __forceinline__ float log1p_uniform_count(long long u, int n) {
    const double s = (double)n * -expm1(-(double)u / (double)n);
    return (float)log1p(s);
}

__device__ __forceinline__ void normalize(const float4 f, const float4 mu, const float4 sd,
                                          float x[kF]) {
    x[0] = (f.x - mu.x) / sd.x;
    x[1] = (f.y - mu.y) / sd.y;
    x[2] = (f.z - mu.z) / sd.z;
    x[3] = (f.w - mu.w) / sd.w;
}

// Forward pass of one sample with the parameters in LDS; fixed summation
// order, so inference is bitwise reproducible.
__device__ __forceinline__ float mlp_forward(const float* P, const float x[kF], float h1[kH]) {
#pragma unroll
    for (int j = 0; j < kH; ++j) {
        float z = P[kOffB1 + j];
#pragma unroll
        for (int f = 0; f < kF; ++f) z = fmaf(P[kOffW1 + j * kF + f], x[f], z);
        h1[j] = relu(z);
    }
    float y = P[kOffB3];
    for (int j = 0; j < kH; ++j) {
        float z = P[kOffB2 + j];
#pragma unroll
        for (int i = 0; i < kH; ++i) z = fmaf(P[kOffW2 + j * kH + i], h1[i], z);
        y = fmaf(P[kOffW3 + j], relu(z), y);
    }
    return y;
}

__global__ __launch_bounds__(kBlock) void k_mlp_predict(const float4* __restrict__ feat, int m,
                                                         const float* __restrict__ Pg, float4 mu,
                                                         float4 sd, const long long* __restrict__ ub,
                                                         int n, const int2* __restrict__ rwin,
                                                         float scale, float* __restrict__ pred) {
    __shared__ float P[kNP];
    for (int q = threadIdx.x; q < kNP; q += kBlock) P[q] = Pg[q];
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) return;
    const long long u = ub[i];
    if (u == 0) {
        pred[i] = 0.f;
        return;
    }
    float x[kF], h1[kH];
    normalize(feat[i], mu, sd, x);
    const float y = mlp_forward(P, x, h1);
    const int ni = rwin ? rwin[i].y : n;  // columns the row can reach
    float p = expm1f(y + log1p_uniform_count(u, ni)) * scale;
    const float hi = (float)min(u, (long long)ni);
    p = p > 0.f ? p : 0.f;  // also maps NaN to 0
    p = p > hi ? hi : p;
    pred[i] = p;
}

// Gradient of the mean squared error in the log-ratio space over one
// minibatch; partial[blockIdx.x][p] = sum over this workgroup's samples.
__global__ __launch_bounds__(kBlock) void k_mlp_grad(const float4* __restrict__ feat,
                                                      const float* __restrict__ target,
                                                      const int* __restrict__ order,
                                                      long long first, long long count,
                                                      float inv_batch, const float* __restrict__ Pg,
                                                      float4 mu, float4 sd,
                                                      float* __restrict__ partial,
                                                      float* __restrict__ loss_part) {
    __shared__ float P[kNP];
    __shared__ float sx[kSamplesPerWG][kF];
    __shared__ float sh1[kSamplesPerWG][kH];
    __shared__ float sh2[kSamplesPerWG][kH];
    __shared__ float sd2[kSamplesPerWG][kH];
    __shared__ float sd1[kSamplesPerWG][kH];
    __shared__ float sg[kSamplesPerWG];
    __shared__ float sloss[kSamplesPerWG];
    for (int q = threadIdx.x; q < kNP; q += kBlock) P[q] = Pg[q];
    __syncthreads();
    const long long s0 = first + (long long)blockIdx.x * kSamplesPerWG;
    if (threadIdx.x < kSamplesPerWG) {
        const int s = threadIdx.x;
        const long long q = s0 + s;
        float x[kF] = {0.f, 0.f, 0.f, 0.f};
        float g = 0.f, l = 0.f;
        float h1[kH];
        if (q < first + count) {
            const int idx = order[q];
            normalize(feat[idx], mu, sd, x);
            const float y = mlp_forward(P, x, h1);
            const float d = y - target[idx];
            g = 2.f * d * inv_batch;
            l = d * d;
        } else {
#pragma unroll
            for (int j = 0; j < kH; ++j) h1[j] = 0.f;
        }
#pragma unroll
        for (int f = 0; f < kF; ++f) sx[s][f] = x[f];
        for (int j = 0; j < kH; ++j) sh1[s][j] = h1[j];
        // second layer again (activations and the output delta)
        for (int j = 0; j < kH; ++j) {
            float z = P[kOffB2 + j];
            for (int i = 0; i < kH; ++i) z = fmaf(P[kOffW2 + j * kH + i], h1[i], z);
            sh2[s][j] = relu(z);
            sd2[s][j] = z > 0.f ? g * P[kOffW3 + j] : 0.f;
        }
        for (int i = 0; i < kH; ++i) {
            float d = 0.f;
            for (int j = 0; j < kH; ++j) d = fmaf(sd2[s][j], P[kOffW2 + j * kH + i], d);
            sd1[s][i] = h1[i] > 0.f ? d : 0.f;
        }
        sg[s] = g;
        sloss[s] = l;
    }
    __syncthreads();
    float* out = partial + (long long)blockIdx.x * kNP;
    for (int p = threadIdx.x; p < kNP; p += kBlock) {
        float acc = 0.f;
        if (p < kOffB1) {
            const int j = p / kF, f = p % kF;
            for (int s = 0; s < kSamplesPerWG; ++s) acc = fmaf(sd1[s][j], sx[s][f], acc);
        } else if (p < kOffW2) {
            const int j = p - kOffB1;
            for (int s = 0; s < kSamplesPerWG; ++s) acc += sd1[s][j];
        } else if (p < kOffB2) {
            const int j = (p - kOffW2) / kH, i = (p - kOffW2) % kH;
            for (int s = 0; s < kSamplesPerWG; ++s) acc = fmaf(sd2[s][j], sh1[s][i], acc);
        } else if (p < kOffW3) {
            const int j = p - kOffB2;
            for (int s = 0; s < kSamplesPerWG; ++s) acc += sd2[s][j];
        } else if (p < kOffB3) {
            const int j = p - kOffW3;
            for (int s = 0; s < kSamplesPerWG; ++s) acc = fmaf(sg[s], sh2[s][j], acc);
        } else {
            for (int s = 0; s < kSamplesPerWG; ++s) acc += sg[s];
        }
        out[p] = acc;
    }
    if (threadIdx.x == 0) {
        float l = 0.f;
        for (int s = 0; s < kSamplesPerWG; ++s) l += sloss[s];
        loss_part[blockIdx.x] = l;
    }
}

// Adam step; the gradient is the sum of the workgroup partials in order.
__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ M, float* __restrict__ Vv,
                           const float* __restrict__ partial, int nwg, float lr, float b1,
                           float b2, float b1t, float b2t) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNP) return;
    float g = 0.f;
    for (int w = 0; w < nwg; ++w) g += partial[(long long)w * kNP + p];
    const float m = b1 * M[p] + (1.f - b1) * g;
    const float v = b2 * Vv[p] + (1.f - b2) * g * g;
    M[p] = m;
    Vv[p] = v;
    const float mh = m / (1.f - b1t), vh = v / (1.f - b2t);
    P[p] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// Samples of one product: features and log-ratio targets of its non-empty
// rows (count from the symbolic pass, crow = C's row offsets).
__global__ void k_mlp_samples(const float4* __restrict__ feat, const long long* __restrict__ ub,
                              const long long* __restrict__ crow, int m, int n,
                              const int2* __restrict__ rwin,
                              const int* __restrict__ pos, float4* __restrict__ sfeat,
                              float* __restrict__ starget) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || ub[i] == 0) return;
    const int q = pos[i];
    sfeat[q] = feat[i];
    const float cnt = (float)(crow[i + 1] - crow[i]);
    starget[q] = log1pf(cnt) - log1p_uniform_count(ub[i], rwin ? rwin[i].y : n);
}

__global__ void k_nonempty_flag(const long long* __restrict__ ub, int m, int* __restrict__ flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) flag[i] = ub[i] > 0 ? 1 : 0;
}

}  // namespace spg

using namespace spg;

struct spg_predictor {
    int device = 0;
    hipStream_t stream = nullptr;  // the null stream
    uint64_t seed = 0;
    float mu[4] = {0, 0, 0, 0};
    float sd[4] = {1, 1, 1, 1};
    bool normalized = false;
    DevBuf P, M, Vv;              // parameters and Adam moments
    DevBuf sfeat, starget;        // training samples
    long long nsamples = 0;
    long long adam_t = 0;
    long long version = 0;  // bumped whenever the weights change
    std::vector<float> host_feat;  // sample features (for the normalization)
    std::string err;
};

namespace {

int pfail(spg_predictor* p, int code, const std::string& m) {
    if (p) p->err = m;
    return code;
}

#define PHIP(call)                                                                  \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) return pfail(p, SPG_EHIP, std::string(#call ": ") +   \
                                                     hipGetErrorString(e_));        \
    } while (0)

int pensure(spg_predictor* p, DevBuf& b, size_t bytes) {
    if (b.cap >= bytes) return SPG_OK;
    void* nb = nullptr;
    if (hipMalloc(&nb, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return pfail(p, SPG_ENOMEM, "predictor allocation");
    }
    if (b.p) {
        PHIP(hipMemcpyAsync(nb, b.p, b.cap, hipMemcpyDeviceToDevice, p->stream));
        PHIP(hipStreamSynchronize(p->stream));
        (void)hipFree(b.p);
    }
    b.p = nb;
    b.cap = bytes;
    return SPG_OK;
}

void init_params(std::vector<float>& w, uint64_t seed) {
    Xoshiro r(seed);
    auto gauss = [&]() {
        double u1 = r.uniform01(), u2 = r.uniform01();
        if (u1 < 1e-300) u1 = 1e-300;
        return (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    };
    w.assign(kNP, 0.f);
    for (int q = 0; q < kH * kF; ++q) w[kOffW1 + q] = gauss() * std::sqrt(2.f / kF);
    for (int q = 0; q < kH * kH; ++q) w[kOffW2 + q] = gauss() * std::sqrt(2.f / kH);
    for (int q = 0; q < kH; ++q) w[kOffW3 + q] = gauss() * std::sqrt(1.f / kH);
    for (int q = 0; q < kH; ++q) w[kOffB1 + q] = 0.01f;
    for (int q = 0; q < kH; ++q) w[kOffB2 + q] = 0.01f;
    w[kOffB3] = -0.5f;
}

}  // namespace

namespace spg {
// Used by the learned path in spg_api.hip.
int predictor_run(spg_context* c, const spg_predictor* p, float scale, float* d_pred) {
    if (c->m <= 0) return SPG_OK;
    const float4 mu = make_float4(p->mu[0], p->mu[1], p->mu[2], p->mu[3]);
    const float4 sd = make_float4(p->sd[0], p->sd[1], p->sd[2], p->sd[3]);
    hipLaunchKernelGGL(k_mlp_predict, dim3((c->m + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       c->stream, (const float4*)c->feat.p, c->m, (const float*)p->P.p, mu, sd,
                       (const long long*)c->ub.p, c->n, (const int2*)c->rwin, scale, d_pred);
    SPG_HIP(hipGetLastError());
    return SPG_OK;
}
int predictor_device(const spg_predictor* p) { return p->device; }
long long predictor_version(const spg_predictor* p) { return p->version; }
}  // namespace spg

extern "C" {

int spgPredictorCreate(spg_predictor** out, int device, uint64_t seed) {
    if (!out) return SPG_EINVAL;
    *out = nullptr;
    int n = 0;
    if (spgDeviceCount(&n) != SPG_OK || device < 0 || device >= n) return SPG_ENODEV;
    spg_predictor* p = new (std::nothrow) spg_predictor();
    if (!p) return SPG_ENOMEM;
    p->device = device;
    p->seed = seed;
    // the predictor's copies and training kernels go to the device's null
    // stream: a stream of its own would take one of the process's few
    // hardware queues (4 by default), and a context's merge streams that
    // end up sharing a queue lose their overlap
    if (hipSetDevice(device) != hipSuccess) {
        delete p;
        return SPG_EHIP;
    }
    std::vector<float> w;
    init_params(w, seed);
    int rc = pensure(p, p->P, sizeof(float) * kNP);
    if (rc == SPG_OK) rc = pensure(p, p->M, sizeof(float) * kNP);
    if (rc == SPG_OK) rc = pensure(p, p->Vv, sizeof(float) * kNP);
    if (rc == SPG_OK) {
        std::vector<float> z(kNP, 0.f);
        if (hipMemcpy(p->P.p, w.data(), sizeof(float) * kNP, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->M.p, z.data(), sizeof(float) * kNP, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->Vv.p, z.data(), sizeof(float) * kNP, hipMemcpyHostToDevice) != hipSuccess)
            rc = SPG_EHIP;
    }
    if (rc != SPG_OK) {
        spgPredictorDestroy(p);
        return rc;
    }
    *out = p;
    return SPG_OK;
}

int spgPredictorDestroy(spg_predictor* p) {
    if (!p) return SPG_EINVAL;
    (void)hipSetDevice(p->device);
    (void)hipStreamSynchronize(p->stream);
    for (DevBuf* b : {&p->P, &p->M, &p->Vv, &p->sfeat, &p->starget})
        if (b->p) (void)hipFree(b->p);
    delete p;
    return SPG_OK;
}

const char* spgPredictorGetLastError(spg_predictor* p) { return p ? p->err.c_str() : ""; }

int spgPredictorNumParams(void) { return kNP; }

int spgPredictorAddSamples(spg_predictor* p, spg_context* c) {
    if (!p || !c) return SPG_EINVAL;
    if (c->stage < 4) return pfail(p, SPG_EORDER, "SymbolicGEMM must run on the context first");
    if (c->device != p->device) return pfail(p, SPG_EINVAL, "context is on another device");
    PHIP(hipSetDevice(p->device));
    PHIP(hipStreamSynchronize(c->stream));
    const int m = c->m;
    // positions of the non-empty rows (exclusive scan of a flag), on c's stream
    int* flag = (int*)scratch_alloc(c, sizeof(int) * ((size_t)m + 1));
    int* pos = (int*)scratch_alloc(c, sizeof(int) * ((size_t)m + 1));
    if (!flag || !pos) {
        scratch_reset(c);
        flag = (int*)scratch_alloc(c, sizeof(int) * ((size_t)m + 1));
        pos = (int*)scratch_alloc(c, sizeof(int) * ((size_t)m + 1));
        if (!flag || !pos) return pfail(p, SPG_ENOMEM, "sample scratch");
    }
    hipLaunchKernelGGL(k_nonempty_flag, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream,
                       (const long long*)c->ub.p, m, flag);
    PHIP(hipGetLastError());
    if (scan_i32(c, flag, pos, m) != SPG_OK) return pfail(p, SPG_EHIP, spgGetLastError(c));
    int added = 0;
    PHIP(hipMemcpyAsync(&added, pos + m, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PHIP(hipStreamSynchronize(c->stream));
    const long long n0 = p->nsamples;
    int rc = pensure(p, p->sfeat, sizeof(float4) * (size_t)(n0 + added + 1));
    if (rc == SPG_OK) rc = pensure(p, p->starget, sizeof(float) * (size_t)(n0 + added + 1));
    if (rc != SPG_OK) return rc;
    hipLaunchKernelGGL(k_mlp_samples, dim3((m + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream,
                       (const float4*)c->feat.p, (const long long*)c->ub.p,
                       (const long long*)c->crow.p, m, c->n, (const int2*)c->rwin, pos, (float4*)p->sfeat.p + n0,
                       (float*)p->starget.p + n0);
    PHIP(hipGetLastError());
    std::vector<float> hf((size_t)added * 4);
    if (added > 0)
        PHIP(hipMemcpyAsync(hf.data(), (float4*)p->sfeat.p + n0, sizeof(float4) * (size_t)added,
                            hipMemcpyDeviceToHost, c->stream));
    PHIP(hipStreamSynchronize(c->stream));
    p->host_feat.insert(p->host_feat.end(), hf.begin(), hf.end());
    p->nsamples += added;
    return SPG_OK;
}

namespace {
// Feature normalization (mean and standard deviation over the samples),
// fixed the first time the network is trained or its gradient is taken,
// unless weights with their own normalization were set.
void normalize_once(spg_predictor* p) {
    if (p->normalized) return;
    const long long N = p->nsamples;
    for (int f = 0; f < 4; ++f) {
        double s = 0, s2 = 0;
        for (long long q = 0; q < N; ++q) {
            double x = p->host_feat[(size_t)q * 4 + f];
            s += x;
            s2 += x * x;
        }
        const double mean = s / N, var = s2 / N - mean * mean;
        p->mu[f] = (float)mean;
        p->sd[f] = (float)(var > 1e-12 ? std::sqrt(var) : 1.0);
    }
    p->normalized = true;
    ++p->version;
}
}  // namespace

int spgPredictorGradient(spg_predictor* p, float* grad, float* loss) {
    if (!p || !grad) return SPG_EINVAL;
    if (p->nsamples <= 0) return pfail(p, SPG_EORDER, "no training samples");
    PHIP(hipSetDevice(p->device));
    normalize_once(p);
    const long long N = p->nsamples;
    const int g = (int)((N + kSamplesPerWG - 1) / kSamplesPerWG);
    DevBuf order, partial, lossb;
    auto cleanup = [&]() {
        for (DevBuf* b : {&order, &partial, &lossb})
            if (b->p) (void)hipFree(b->p);
    };
    int rc = pensure(p, order, sizeof(int) * (size_t)N);
    if (rc == SPG_OK) rc = pensure(p, partial, sizeof(float) * (size_t)g * kNP);
    if (rc == SPG_OK) rc = pensure(p, lossb, sizeof(float) * (size_t)g);
    if (rc != SPG_OK) {
        cleanup();
        return rc;
    }
    std::vector<int> ident((size_t)N);
    for (long long q = 0; q < N; ++q) ident[(size_t)q] = (int)q;
    std::vector<float> hp((size_t)g * kNP), hl((size_t)g);
    const float4 mu = make_float4(p->mu[0], p->mu[1], p->mu[2], p->mu[3]);
    const float4 sd = make_float4(p->sd[0], p->sd[1], p->sd[2], p->sd[3]);
    bool ok = hipMemcpyAsync(order.p, ident.data(), sizeof(int) * (size_t)N,
                             hipMemcpyHostToDevice, p->stream) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_mlp_grad, dim3(g), dim3(kBlock), 0, p->stream,
                           (const float4*)p->sfeat.p, (const float*)p->starget.p,
                           (const int*)order.p, 0LL, N, 1.f / (float)N, (const float*)p->P.p, mu,
                           sd, (float*)partial.p, (float*)lossb.p);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(hp.data(), partial.p, sizeof(float) * hp.size(),
                            hipMemcpyDeviceToHost, p->stream) == hipSuccess &&
             hipMemcpyAsync(hl.data(), lossb.p, sizeof(float) * hl.size(), hipMemcpyDeviceToHost,
                            p->stream) == hipSuccess &&
             hipStreamSynchronize(p->stream) == hipSuccess;
    }
    cleanup();
    if (!ok) return pfail(p, SPG_EHIP, "gradient kernels failed");
    for (int q = 0; q < kNP; ++q) {
        double acc = 0.0;
        for (int w = 0; w < g; ++w) acc += hp[(size_t)w * kNP + q];
        grad[q] = (float)acc;
    }
    if (loss) {
        double l = 0.0;
        for (int w = 0; w < g; ++w) l += hl[(size_t)w];
        *loss = (float)(l / (double)N);
    }
    return SPG_OK;
}

int spgPredictorTrain(spg_predictor* p, int epochs, int batch, float lr, float* final_loss) {
    if (!p || epochs < 0 || batch < kSamplesPerWG || lr <= 0.f) return SPG_EINVAL;
    if (p->nsamples <= 0) return pfail(p, SPG_EORDER, "no training samples");
    PHIP(hipSetDevice(p->device));
    const long long N = p->nsamples;
    normalize_once(p);
    batch = (batch / kSamplesPerWG) * kSamplesPerWG;
    const int nwg = batch / kSamplesPerWG;
    DevBuf order, partial, lossb;
    auto cleanup = [&]() {
        for (DevBuf* b : {&order, &partial, &lossb})
            if (b->p) (void)hipFree(b->p);
    };
    int rc = pensure(p, order, sizeof(int) * (size_t)N);
    if (rc == SPG_OK) rc = pensure(p, partial, sizeof(float) * (size_t)nwg * kNP);
    if (rc == SPG_OK) rc = pensure(p, lossb, sizeof(float) * (size_t)nwg);
    if (rc != SPG_OK) {
        cleanup();
        return rc;
    }
    std::vector<int> perm((size_t)N);
    Xoshiro r(p->seed ^ 0x5bd1e995ull);
    const float4 mu = make_float4(p->mu[0], p->mu[1], p->mu[2], p->mu[3]);
    const float4 sd = make_float4(p->sd[0], p->sd[1], p->sd[2], p->sd[3]);
    const float b1 = 0.9f, b2 = 0.999f;
    double last = 0.0;
    std::vector<float> hl((size_t)nwg);
    for (int ep = 0; ep < epochs; ++ep) {
        for (long long q = 0; q < N; ++q) perm[(size_t)q] = (int)q;
        for (long long q = N - 1; q > 0; --q) std::swap(perm[(size_t)q], perm[(size_t)r.below((uint64_t)q + 1)]);
        if (hipMemcpyAsync(order.p, perm.data(), sizeof(int) * (size_t)N, hipMemcpyHostToDevice,
                           p->stream) != hipSuccess) {
            cleanup();
            return pfail(p, SPG_EHIP, "order upload");
        }
        double ep_loss = 0.0;
        for (long long first = 0; first < N; first += batch) {
            const long long cnt = std::min<long long>(batch, N - first);
            const int g = (int)((cnt + kSamplesPerWG - 1) / kSamplesPerWG);
            hipLaunchKernelGGL(k_mlp_grad, dim3(g), dim3(kBlock), 0, p->stream,
                               (const float4*)p->sfeat.p, (const float*)p->starget.p,
                               (const int*)order.p, first, cnt, 1.f / (float)cnt,
                               (const float*)p->P.p, mu, sd, (float*)partial.p, (float*)lossb.p);
            ++p->adam_t;
            const float b1t = std::pow(b1, (float)p->adam_t), b2t = std::pow(b2, (float)p->adam_t);
            hipLaunchKernelGGL(k_mlp_adam, dim3((kNP + kBlock - 1) / kBlock), dim3(kBlock), 0,
                               p->stream, (float*)p->P.p, (float*)p->M.p, (float*)p->Vv.p,
                               (const float*)partial.p, g, lr, b1, b2, b1t, b2t);
            if (ep == epochs - 1) {
                if (hipMemcpyAsync(hl.data(), lossb.p, sizeof(float) * g, hipMemcpyDeviceToHost,
                                   p->stream) != hipSuccess ||
                    hipStreamSynchronize(p->stream) != hipSuccess) {
                    cleanup();
                    return pfail(p, SPG_EHIP, "loss readback");
                }
                for (int w = 0; w < g; ++w) ep_loss += hl[(size_t)w];
            }
        }
        if (ep == epochs - 1) last = ep_loss / (double)N;
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess || hipGetLastError() != hipSuccess) {
        cleanup();
        return pfail(p, SPG_EHIP, "training kernels failed");
    }
    cleanup();
    ++p->version;
    if (final_loss) *final_loss = (float)last;
    return SPG_OK;
}

int spgPredictorGetWeights(spg_predictor* p, float* w, float* norm8) {
    if (!p || !w) return SPG_EINVAL;
    PHIP(hipSetDevice(p->device));
    PHIP(hipMemcpy(w, p->P.p, sizeof(float) * kNP, hipMemcpyDeviceToHost));
    if (norm8) {
        for (int f = 0; f < 4; ++f) {
            norm8[f] = p->mu[f];
            norm8[4 + f] = p->sd[f];
        }
    }
    return SPG_OK;
}

int spgPredictorSetWeights(spg_predictor* p, const float* w, const float* norm8) {
    if (!p || !w || !norm8) return SPG_EINVAL;
    PHIP(hipSetDevice(p->device));
    PHIP(hipMemcpy(p->P.p, w, sizeof(float) * kNP, hipMemcpyHostToDevice));
    ++p->version;
    for (int f = 0; f < 4; ++f) {
        p->mu[f] = norm8[f];
        p->sd[f] = norm8[4 + f] > 0.f ? norm8[4 + f] : 1.f;
    }
    p->normalized = true;
    return SPG_OK;
}

// File format: 8-byte magic "SPGMLP01", int32 parameter count, 8 floats of
// normalization (mean, sd), then the parameters (little-endian FP32).
int spgPredictorSave(spg_predictor* p, const char* path) {
    if (!p || !path) return SPG_EINVAL;
    std::vector<float> w(kNP);
    float norm[8];
    int rc = spgPredictorGetWeights(p, w.data(), norm);
    if (rc != SPG_OK) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return pfail(p, SPG_EIO, "cannot open predictor file for writing");
    const int32_t np = kNP;
    bool ok = fwrite("SPGMLP01", 1, 8, f) == 8 && fwrite(&np, 4, 1, f) == 1 &&
              fwrite(norm, 4, 8, f) == 8 && fwrite(w.data(), 4, kNP, f) == (size_t)kNP;
    ok = (fclose(f) == 0) && ok;
    return ok ? SPG_OK : pfail(p, SPG_EIO, "predictor write failed");
}

int spgPredictorLoad(spg_predictor* p, const char* path) {
    if (!p || !path) return SPG_EINVAL;
    FILE* f = fopen(path, "rb");
    if (!f) return pfail(p, SPG_EIO, "cannot open predictor file");
    char magic[8];
    int32_t np = 0;
    float norm[8];
    std::vector<float> w(kNP);
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "SPGMLP01", 8) == 0 &&
              fread(&np, 4, 1, f) == 1 && np == kNP && fread(norm, 4, 8, f) == 8 &&
              fread(w.data(), 4, kNP, f) == (size_t)kNP;
    fclose(f);
    if (!ok) return pfail(p, SPG_EIO, "not a predictor file of this version");
    return spgPredictorSetWeights(p, w.data(), norm);
}

}  // extern "C"
