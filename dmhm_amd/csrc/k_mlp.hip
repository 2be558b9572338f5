// gfx950 kernels of the rank model: a 3-layer perceptron (NF inputs, two
// hidden layers of 64 ReLU units, one linear output) that predicts the
// numerical rank of an admissible block from its features. Inference for a
// whole product's blocks, and training (forward + backward of mean squared
// error over mini-batches, Adam update), all in FP32.
//
// Parameter layout (one float array, MLP_PARAMS(NF) entries):
//   W1 [NF x 64] (row f: weights from input f), b1 [64],
//   W2 [64 x 64] (row i: weights from hidden unit i), b2 [64], w3 [64], b3 [1].
#include "kcommon.hpp"

namespace dmhm {

constexpr int MLP_H = 64;
__host__ __device__ constexpr int mlp_params(int nf) { return nf * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H + 1; }

struct MlpView {
    const float *W1, *b1, *W2, *b2, *w3, *b3;
};
__device__ inline MlpView mlp_view(const float* p, int nf)
{
    MlpView v;
    v.W1 = p; v.b1 = v.W1 + nf * MLP_H; v.W2 = v.b1 + MLP_H; v.b2 = v.W2 + MLP_H * MLP_H;
    v.w3 = v.b2 + MLP_H; v.b3 = v.w3 + MLP_H;
    return v;
}

// One thread per sample; the parameters are staged in LDS once per workgroup.
// y[i] = model output; widths[i] (if given) = clamp(ceil(y + margin) + extra,
// wmin, wmax).
__global__ void __launch_bounds__(256) mlp_infer_kernel(const float* __restrict__ params, int nf,
                                                        const float* __restrict__ X, int n, float* y,
                                                        int* widths, float margin, int extra, int wmin,
                                                        int wmax)
{
    extern __shared__ float sp[];
    const int np = mlp_params(nf);
    for (int e = threadIdx.x; e < np; e += blockDim.x) sp[e] = params[e];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const MlpView v = mlp_view(sp, nf);
    float h1[MLP_H];
#pragma unroll
    for (int j = 0; j < MLP_H; ++j) h1[j] = v.b1[j];
    for (int f = 0; f < nf; ++f) {
        const float x = X[(int64_t)i * nf + f];
#pragma unroll
        for (int j = 0; j < MLP_H; ++j) h1[j] += x * v.W1[f * MLP_H + j];
    }
    float out = v.b3[0];
    for (int j = 0; j < MLP_H; ++j) {
        float h2 = v.b2[j];
#pragma unroll
        for (int k = 0; k < MLP_H; ++k) h2 += fmaxf(h1[k], 0.0f) * v.W2[k * MLP_H + j];
        out += fmaxf(h2, 0.0f) * v.w3[j];
    }
    if (y) y[i] = out;
    if (widths) {
        int w = (int)ceilf(out + margin) + extra;
        widths[i] = min(max(w, wmin), wmax);
    }
}

// Gradient of L = (1/n) sum_i (y_i - t_i)^2 over samples [0, n): one
// workgroup of 64 threads per 64 samples (thread = sample) keeps the
// activations in LDS, then reduces the weight gradients over its samples and
// adds them to `grads` (atomically, one add per parameter per workgroup).
// loss (if given) accumulates sum (y - t)^2 / n.
__global__ void __launch_bounds__(64) mlp_grad_kernel(const float* __restrict__ params, int nf,
                                                      const float* __restrict__ X, const float* __restrict__ t,
                                                      int n, float* grads, float* loss)
{
    constexpr int B = 64;
    __shared__ float Xs[B][16];
    __shared__ float H1[B][MLP_H + 1];     // pre-activations, then activations
    __shared__ float H2[B][MLP_H + 1];
    __shared__ float G1[B][MLP_H + 1];     // dL/d pre-activation 1
    __shared__ float G2[B][MLP_H + 1];     // dL/d pre-activation 2
    __shared__ float dy[B];
    const MlpView v = mlp_view(params, nf);
    const int tid = threadIdx.x;
    const int i = blockIdx.x * B + tid;
    const bool live = i < n;
    const float scale = 2.0f / (float)n;
    for (int f = 0; f < nf; ++f) Xs[tid][f] = live ? X[(int64_t)i * nf + f] : 0.0f;
    // Forward.
    for (int j = 0; j < MLP_H; ++j) {
        float a = v.b1[j];
        for (int f = 0; f < nf; ++f) a += Xs[tid][f] * v.W1[f * MLP_H + j];
        H1[tid][j] = fmaxf(a, 0.0f);
    }
    float out = v.b3[0];
    for (int j = 0; j < MLP_H; ++j) {
        float a = v.b2[j];
        for (int k = 0; k < MLP_H; ++k) a += H1[tid][k] * v.W2[k * MLP_H + j];
        H2[tid][j] = fmaxf(a, 0.0f);
        out += H2[tid][j] * v.w3[j];
    }
    const float r = live ? out - t[i] : 0.0f;
    dy[tid] = scale * r;
    if (loss) {
        float l = r * r / (float)n;
        l = wave_sum(l);
        if (tid == 0) atomicAdd(loss, l);
    }
    // Backward through the samples' own activations.
    for (int j = 0; j < MLP_H; ++j) G2[tid][j] = H2[tid][j] > 0.0f ? dy[tid] * v.w3[j] : 0.0f;
    for (int k = 0; k < MLP_H; ++k) {
        float a = 0.0f;
        if (H1[tid][k] > 0.0f)
            for (int j = 0; j < MLP_H; ++j) a += G2[tid][j] * v.W2[k * MLP_H + j];
        G1[tid][k] = a;
    }
    __syncthreads();
    // Parameter gradients, reduced over the workgroup's samples.
    float* gW1 = grads;
    float* gb1 = gW1 + nf * MLP_H;
    float* gW2 = gb1 + MLP_H;
    float* gb2 = gW2 + MLP_H * MLP_H;
    float* gw3 = gb2 + MLP_H;
    float* gb3 = gw3 + MLP_H;
    for (int e = tid; e < MLP_H * MLP_H; e += B) {       // W2[k][j] += H1[b][k] G2[b][j]
        const int k = e / MLP_H, j = e % MLP_H;
        float a = 0.0f;
        for (int b2 = 0; b2 < B; ++b2) a += H1[b2][k] * G2[b2][j];
        atomicAdd(&gW2[e], a);
    }
    for (int e = tid; e < nf * MLP_H; e += B) {          // W1[f][j] += X[b][f] G1[b][j]
        const int f = e / MLP_H, j = e % MLP_H;
        float a = 0.0f;
        for (int b2 = 0; b2 < B; ++b2) a += Xs[b2][f] * G1[b2][j];
        atomicAdd(&gW1[e], a);
    }
    {
        const int j = tid;                              // B == MLP_H: one unit per thread
        float s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        for (int b2 = 0; b2 < B; ++b2) {
            s1 += G1[b2][j];
            s2 += G2[b2][j];
            s3 += H2[b2][j] * dy[b2];
        }
        atomicAdd(&gb1[j], s1);
        atomicAdd(&gb2[j], s2);
        atomicAdd(&gw3[j], s3);
    }
    if (tid == 0) {
        float s = 0.0f;
        for (int b2 = 0; b2 < B; ++b2) s += dy[b2];
        atomicAdd(gb3, s);
    }
}

// Adam step (Kingma & Ba) on every parameter; clears the gradient.
__global__ void mlp_adam_kernel(float* p, float* g, float* m, float* v, int np, float lr, float b1, float b2,
                                float eps, float c1, float c2)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= np) return;
    const float gr = g[e];
    const float mm = b1 * m[e] + (1.0f - b1) * gr;
    const float vv = b2 * v[e] + (1.0f - b2) * gr * gr;
    m[e] = mm; v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    g[e] = 0.0f;
}

void launch_mlp_infer(const float* params, int nf, const float* X, int n, float* y, int* widths, float margin,
                      int extra, int wmin, int wmax, hipStream_t s)
{
    if (n <= 0) return;
    if (nf < 1 || nf > 16) throw std::runtime_error("rank model takes 1 to 16 features");
    const size_t lds = (size_t)mlp_params(nf) * sizeof(float);
    mlp_infer_kernel<<<(n + 255) / 256, 256, lds, s>>>(params, nf, X, n, y, widths, margin, extra, wmin, wmax);
    HIP_LAUNCH_CHECK();
}

void launch_mlp_grad(const float* params, int nf, const float* X, const float* t, int n, float* grads,
                     float* loss, hipStream_t s)
{
    if (n <= 0) return;
    if (nf < 1 || nf > 16) throw std::runtime_error("rank model takes 1 to 16 features");
    mlp_grad_kernel<<<(n + 63) / 64, 64, 0, s>>>(params, nf, X, t, n, grads, loss);
    HIP_LAUNCH_CHECK();
}

void launch_mlp_adam(float* p, float* g, float* m, float* v, int np, float lr, int step, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    mlp_adam_kernel<<<(np + 255) / 256, 256, 0, s>>>(p, g, m, v, np, lr, b1, b2, 1e-8f, c1, c2);
    HIP_LAUNCH_CHECK();
}

int mlp_param_count(int nf) { return mlp_params(nf); }

// Block features for the rank model: the static features of each admissible
// block (nstat per block) followed by its "previous rank" (the rank of the
// same block in a given rank array, scaled by 1/maxRank; 0 without one).
__global__ void rank_features_kernel(const float* __restrict__ stat, int nstat, const int* __restrict__ prev,
                                     float invMax, float* __restrict__ out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int f = 0; f < nstat; ++f) out[(int64_t)i * (nstat + 1) + f] = stat[(int64_t)i * nstat + f];
    out[(int64_t)i * (nstat + 1) + nstat] = prev ? (float)prev[i] * invMax : 0.0f;
}
void launch_rank_features(const float* stat, int nstat, const int* prev, float invMax, float* out, int n,
                          hipStream_t s)
{
    if (n <= 0) return;
    rank_features_kernel<<<(n + 255) / 256, 256, 0, s>>>(stat, nstat, prev, invMax, out, n);
    HIP_LAUNCH_CHECK();
}

// Widths of flagged blocks raised to `wmax` (the full sketch), flags cleared;
// count (if given) receives the number of flagged blocks and wsum the sum of
// the widths before widening.
__global__ void widen_flagged_kernel(int* __restrict__ widths, int* __restrict__ flags, int n, int wmax,
                                     int* count, long long* wsum)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int f = 0, w = 0;
    if (i < n) {
        f = flags[i];
        w = widths[i];
        if (f) widths[i] = wmax;
        flags[i] = 0;
    }
    // one atomic per wave
    const unsigned long long fb = __ballot(f != 0);
    int ws = w;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) ws += __shfl_xor(ws, off, 64);
    if ((threadIdx.x & 63) == 0) {
        if (count) atomicAdd(count, (int)__popcll(fb));
        if (wsum) atomicAdd((unsigned long long*)wsum, (unsigned long long)ws);
    }
}
void launch_widen_flagged(int* widths, int* flags, int n, int wmax, int* count, long long* wsum, hipStream_t s)
{
    if (n <= 0) return;
    widen_flagged_kernel<<<(n + 255) / 256, 256, 0, s>>>(widths, flags, n, wmax, count, wsum);
    HIP_LAUNCH_CHECK();
}

} // namespace dmhm
