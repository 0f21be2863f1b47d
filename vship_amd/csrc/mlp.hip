// mlp.hip - the score surrogate's network on gfx950: forward, backward and
// Adam update of a 55 -> 128 -> 128 -> 2 GELU MLP (mean and log-variance of
// the score) with BF16 weights and FP32 accumulation.
//
// Training is data parallel over workgroups: each workgroup runs forward and
// backward for 32 samples with every activation in LDS and writes its
// gradient sums to a per-workgroup partial row; one update kernel then sums
// the partials in a fixed order (deterministic), applies Adam to the FP32
// master weights and re-quantises the BF16 copies that the forward and
// backward passes read (forward layout: input-major W^T, so consecutive
// threads read consecutive outputs' weights; backward layout: row-major W for
// the input gradients). Biases stay FP32.
//
// Loss per sample (normalised target y, outputs mu, lv):
//   0.5 * (lv + (y - mu)^2 * exp(-lv))   (Gaussian negative log-likelihood)
// averaged over the step's samples.
#include "mlp.hpp"

namespace vship {

namespace {

constexpr int kS = kMlpSamplesPerGroup;  // samples per workgroup
constexpr int kT = kMlpThreads;
static_assert(kT == 2 * kMlpHidden, "thread mapping assumes 2 threads per hidden unit");
static_assert(kS % 2 == 0, "two sample halves per hidden unit");
constexpr int kXPad = kMlpInputs + 1;  // LDS row pitch of the inputs

__device__ inline float bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ inline float gelu_grad(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

struct Lds {
    float x[kS][kXPad];
    float z1[kS][kMlpHidden];
    float h1[kS][kMlpHidden];
    float z2[kS][kMlpHidden];   // backward: dz2
    float h2[kS][kMlpHidden];   // backward: dz1
    float out[kS][kMlpOutputs];
    float d[kS][kMlpOutputs];   // dL/d(mu, lv)
    float y[kS];
    float loss[kS];
};

// Normalised inputs of the workgroup's samples (zeros past n).
__device__ inline void load_inputs(const MlpDevice& m, const float* X, const int* idx, int n, int first, Lds& s)
{
    for (int e = threadIdx.x; e < kS * kMlpInputs; e += kT) {
        const int r = e / kMlpInputs, i = e % kMlpInputs;
        const int g = first + r;
        float v = 0.0f;
        if (g < n) {
            const int64_t row = idx ? idx[g] : g;
            v = (X[row * kMlpInputs + i] - m.xmean[i]) * m.xinv[i];
        }
        s.x[r][i] = v;
    }
}

// Hidden layer: z[s][o] = b[o] + sum_i W[o][i] in[s][i], h = gelu(z); thread
// owns unit o = t % 128 for samples t / 128 + 2k.
template <int K, int IN_PITCH>
__device__ inline void hidden_layer(const uint16_t* __restrict__ wt, const float* __restrict__ b,
                                    const float (*in)[IN_PITCH], float (*z)[kMlpHidden], float (*h)[kMlpHidden])
{
    const int o = threadIdx.x & (kMlpHidden - 1), s0 = threadIdx.x >> 7;
    float acc[kS / 2];
    const float bias = b[o];
#pragma unroll
    for (int k = 0; k < kS / 2; k++) acc[k] = bias;
    for (int i = 0; i < K; i++) {
        const float w = bf16f(wt[i * kMlpHidden + o]);
#pragma unroll
        for (int k = 0; k < kS / 2; k++) acc[k] = fmaf(w, in[s0 + 2 * k][i], acc[k]);
    }
#pragma unroll
    for (int k = 0; k < kS / 2; k++) {
        if (z) z[s0 + 2 * k][o] = acc[k];
        h[s0 + 2 * k][o] = gelu(acc[k]);
    }
}

__device__ inline void forward(const MlpDevice& m, Lds& s)
{
    hidden_layer<kMlpInputs, kXPad>(m.w1t, m.master + kMlpOffB1, s.x, s.z1, s.h1);
    __syncthreads();
    hidden_layer<kMlpHidden, kMlpHidden>(m.w2t, m.master + kMlpOffB2, s.h1, s.z2, s.h2);
    __syncthreads();
    if (threadIdx.x < kS * kMlpOutputs) {
        const int r = threadIdx.x >> 1, j = threadIdx.x & 1;
        float acc = m.master[kMlpOffB3 + j];
        for (int i = 0; i < kMlpHidden; i++) acc = fmaf(bf16f(m.w3t[i * kMlpOutputs + j]), s.h2[r][i], acc);
        s.out[r][j] = acc;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kT) void mlp_grad_kernel(MlpDevice m, const float* __restrict__ X,
                                                      const float* __restrict__ Y, const int* __restrict__ idx,
                                                      int n, float grad_scale, float* __restrict__ partial,
                                                      float* __restrict__ loss_partial)
{
    __shared__ Lds s;
    const int t = threadIdx.x;
    const int first = blockIdx.x * kS;
    load_inputs(m, X, idx, n, first, s);
    if (t < kS) {
        const int g = first + t;
        s.y[t] = g < n ? (Y[idx ? idx[g] : g] - m.ymean) * m.yinv : 0.0f;
    }
    __syncthreads();
    forward(m, s);

    // loss and output gradients
    if (t < kS) {
        const bool valid = first + t < n;
        const float mu = s.out[t][0], lv = s.out[t][1];
        const float r = s.y[t] - mu;
        const float e = expf(-lv);
        s.loss[t] = valid ? 0.5f * (lv + r * r * e) : 0.0f;
        s.d[t][0] = valid ? -r * e * grad_scale : 0.0f;
        s.d[t][1] = valid ? 0.5f * (1.0f - r * r * e) * grad_scale : 0.0f;
    }
    __syncthreads();
    float* g = partial + (int64_t)blockIdx.x * kMlpParams;
    if (t == 0) {
        float l = 0.0f;
        for (int r = 0; r < kS; r++) l += s.loss[r];
        loss_partial[blockIdx.x] = l;
    }
    {
        // dW3[j][i] = sum_s d[s][j] h2[s][i]; db3
        const int j = t >> 7, i = t & (kMlpHidden - 1);
        float acc = 0.0f;
        for (int r = 0; r < kS; r++) acc = fmaf(s.d[r][j], s.h2[r][i], acc);
        g[kMlpOffW3 + j * kMlpHidden + i] = acc;
        if (t < kMlpOutputs) {
            float b = 0.0f;
            for (int r = 0; r < kS; r++) b += s.d[r][t];
            g[kMlpOffB3 + t] = b;
        }
    }
    {
        // dz2[s][i] = (sum_j d[s][j] W3[j][i]) * gelu'(z2[s][i]), into z2
        const int i = t & (kMlpHidden - 1), s0 = t >> 7;
        const float w0 = bf16f(m.w3[i]), w1 = bf16f(m.w3[kMlpHidden + i]);
#pragma unroll
        for (int k = 0; k < kS / 2; k++) {
            const int r = s0 + 2 * k;
            s.z2[r][i] = fmaf(s.d[r][0], w0, s.d[r][1] * w1) * gelu_grad(s.z2[r][i]);
        }
    }
    __syncthreads();
    {
        // dW2[o][i] = sum_s dz2[s][o] h1[s][i]; db2
        const int i = t & (kMlpHidden - 1), o0 = t >> 7;
        float hcol[kS];
#pragma unroll
        for (int r = 0; r < kS; r++) hcol[r] = s.h1[r][i];
        for (int o = o0; o < kMlpHidden; o += 2) {
            float acc = 0.0f;
#pragma unroll
            for (int r = 0; r < kS; r++) acc = fmaf(s.z2[r][o], hcol[r], acc);
            g[kMlpOffW2 + o * kMlpHidden + i] = acc;
        }
        if (t < kMlpHidden) {
            float b = 0.0f;
            for (int r = 0; r < kS; r++) b += s.z2[r][t];
            g[kMlpOffB2 + t] = b;
        }
    }
    {
        // dz1[s][i] = (sum_o dz2[s][o] W2[o][i]) * gelu'(z1[s][i]), into h2
        const int i = t & (kMlpHidden - 1), s0 = t >> 7;
        float acc[kS / 2];
#pragma unroll
        for (int k = 0; k < kS / 2; k++) acc[k] = 0.0f;
        for (int o = 0; o < kMlpHidden; o++) {
            const float w = bf16f(m.w2[o * kMlpHidden + i]);
#pragma unroll
            for (int k = 0; k < kS / 2; k++) acc[k] = fmaf(s.z2[s0 + 2 * k][o], w, acc[k]);
        }
#pragma unroll
        for (int k = 0; k < kS / 2; k++) {
            const int r = s0 + 2 * k;
            s.h2[r][i] = acc[k] * gelu_grad(s.z1[r][i]);
        }
    }
    __syncthreads();
    {
        // dW1[o][i] = sum_s dz1[s][o] x[s][i]; db1
        const int i = t & 63, o0 = t >> 6;
        if (i < kMlpInputs) {
            float xcol[kS];
#pragma unroll
            for (int r = 0; r < kS; r++) xcol[r] = s.x[r][i];
            for (int o = o0; o < kMlpHidden; o += 4) {
                float acc = 0.0f;
#pragma unroll
                for (int r = 0; r < kS; r++) acc = fmaf(s.h2[r][o], xcol[r], acc);
                g[kMlpOffW1 + o * kMlpInputs + i] = acc;
            }
        }
        if (t < kMlpHidden) {
            float b = 0.0f;
            for (int r = 0; r < kS; r++) b += s.h2[r][t];
            g[kMlpOffB1 + t] = b;
        }
    }
}

__device__ inline uint16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                             // round to nearest even
    return (uint16_t)(u >> 16);
}

// BF16 copies of master weight p in the forward and backward layouts.
__device__ inline void write_bf16(const MlpDevice& m, int p, float w)
{
    const uint16_t h = to_bf16(w);
    if (p < kMlpOffB1) {
        const int o = p / kMlpInputs, i = p % kMlpInputs;
        m.w1t[i * kMlpHidden + o] = h;
    } else if (p >= kMlpOffW2 && p < kMlpOffB2) {
        const int q = p - kMlpOffW2, o = q / kMlpHidden, i = q % kMlpHidden;
        m.w2t[i * kMlpHidden + o] = h;
        m.w2[q] = h;
    } else if (p >= kMlpOffW3 && p < kMlpOffB3) {
        const int q = p - kMlpOffW3, j = q / kMlpHidden, i = q % kMlpHidden;
        m.w3t[i * kMlpOutputs + j] = h;
        m.w3[q] = h;
    }
}

// Sum of the workgroup partials in workgroup order; then either the gradient
// is written out (grad != nullptr) or Adam updates the master weight.
__global__ __launch_bounds__(256) void mlp_update_kernel(MlpDevice m, const float* __restrict__ partial, int nparts,
                                                         MlpAdam adam, float* __restrict__ grad,
                                                         const float* __restrict__ loss_partial,
                                                         float* __restrict__ loss_out)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0 && loss_out) {
        float l = 0.0f;
        for (int w = 0; w < nparts; w++) l += loss_partial[w];
        *loss_out = l;
    }
    if (p >= kMlpParams) return;
    float gsum = 0.0f;
    for (int w = 0; w < nparts; w++) gsum += partial[(int64_t)w * kMlpParams + p];
    if (grad) {
        grad[p] = gsum;
        return;
    }
    const float mo = adam.beta1 * m.adam_m[p] + (1.0f - adam.beta1) * gsum;
    const float ve = adam.beta2 * m.adam_v[p] + (1.0f - adam.beta2) * gsum * gsum;
    m.adam_m[p] = mo;
    m.adam_v[p] = ve;
    const float mhat = mo * adam.corr1, vhat = ve * adam.corr2;
    float w = m.master[p];
    w -= adam.lr * (mhat / (sqrtf(vhat) + adam.eps) + adam.weight_decay * w);
    m.master[p] = w;
    write_bf16(m, p, w);
}

__global__ __launch_bounds__(256) void mlp_requant_kernel(MlpDevice m)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < kMlpParams) write_bf16(m, p, m.master[p]);
}

// Inference: mean and standard deviation of the score in score units.
__global__ __launch_bounds__(kT) void mlp_forward_kernel(MlpDevice m, const float* __restrict__ X, int n,
                                                         float* __restrict__ mean, float* __restrict__ sigma)
{
    __shared__ Lds s;
    const int first = blockIdx.x * kS;
    load_inputs(m, X, nullptr, n, first, s);
    __syncthreads();
    forward(m, s);
    const int t = threadIdx.x;
    if (t < kS && first + t < n) {
        mean[first + t] = s.out[t][0] * m.ystd + m.ymean;
        sigma[first + t] = expf(0.5f * s.out[t][1]) * m.ystd;
    }
}

}  // namespace

int mlp_groups(int n) { return (n + kS - 1) / kS; }

hipError_t launch_mlp_grad(const MlpDevice& m, const float* X, const float* Y, const int* idx, int n, float grad_scale,
                           float* partial, float* loss_partial, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(mlp_groups(n)), dim3(kT), 0, stream, m, X, Y, idx, n, grad_scale,
                       partial, loss_partial);
    return hipGetLastError();
}

hipError_t launch_mlp_update(const MlpDevice& m, const float* partial, int nparts, const MlpAdam& adam, float* grad,
                             const float* loss_partial, float* loss_out, hipStream_t stream)
{
    hipLaunchKernelGGL(mlp_update_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, m, partial, nparts,
                       adam, grad, loss_partial, loss_out);
    return hipGetLastError();
}

hipError_t launch_mlp_requant(const MlpDevice& m, hipStream_t stream)
{
    hipLaunchKernelGGL(mlp_requant_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, stream, m);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const MlpDevice& m, const float* X, int n, float* mean, float* sigma,
                              hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(mlp_groups(n)), dim3(kT), 0, stream, m, X, n, mean, sigma);
    return hipGetLastError();
}

}  // namespace vship
