// mlp.hip - the score surrogate's network on gfx950: forward, backward and
// Adam update of a 55 -> 128 -> 128 -> 2 GELU MLP (mean and log-variance of
// the score) with BF16 weights and FP32 accumulation.
//
// Training is data parallel over workgroups: each workgroup runs forward and
// backward for 32 samples with every activation in LDS and writes its
// gradient sums to a per-workgroup partial row; one update kernel then sums
// the partials in a fixed order (deterministic), applies Adam to the FP32
// master weights and re-quantises the BF16 copies that the forward and
// backward passes read. Biases stay FP32.
//
// The layer products run on v_mfma_f32_16x16x32_bf16 (16 x 16 output tiles,
// 32-deep slices, FP32 accumulation). The weights are BF16 already. In the
// products with a weight matrix (both hidden layers, and the input gradient
// of layer 2) the FP32 activations enter as three BF16 terms hi + mid + lo
// (each the round-to-nearest of what the previous ones leave: about 25
// significant bits, the FP32 value), three MFMAs per tile; in the weight
// gradients, products of two FP32 quantities, each enters as hi + lo (about
// 17 bits) and the tile is hi hi + lo hi + hi lo, three MFMAs. The two-unit
// output layer and its gradients stay on the VALU.
//
// Loss per sample (normalised target y, outputs mu, lv):
//   0.5 * (lv + (y - mu)^2 * exp(-lv))   (Gaussian negative log-likelihood)
// averaged over the step's samples.
#include "mlp.hpp"

namespace vship {

namespace {

constexpr int kS = kMlpSamplesPerGroup;  // samples per workgroup
constexpr int kT = kMlpThreads;
static_assert(kT == 2 * kMlpHidden, "the scalar passes map 2 threads to a hidden unit");
static_assert(kT == 256 && kS == 32 && kMlpHidden == 128,
              "matrix-core products: 4 waves of two 16-unit blocks, two 16-sample halves");
static_assert(kMlpInputsPadded % 32 == 0 && kMlpInputsPadded >= kMlpInputs, "32-deep slices of the inputs");
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

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Eight FP32 values as BF16 hi + lo.
__device__ inline void split8(const float (&v)[8], bf16x8& hi, bf16x8& lo)
{
#pragma unroll
    for (int j = 0; j < 8; j++) {
        hi[j] = (__bf16)v[j];
        lo[j] = (__bf16)(v[j] - (float)hi[j]);
    }
}

// Eight FP32 values as BF16 hi + mid + lo.
__device__ inline void split8x3(const float (&v)[8], bf16x8& hi, bf16x8& mid, bf16x8& lo)
{
#pragma unroll
    for (int j = 0; j < 8; j++) {
        hi[j] = (__bf16)v[j];
        const float r = v[j] - (float)hi[j];
        mid[j] = (__bf16)r;
        lo[j] = (__bf16)(r - (float)mid[j]);
    }
}

// Eight consecutive BF16 weights (16-byte aligned).
__device__ inline bf16x8 load_w8(const uint16_t* p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ inline floatx4 mfma(const bf16x8& a, const bf16x8& b, floatx4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// C[s][n] = sum_k in[s][k] W[n][k] for the 32 samples and the wave's two
// 16-unit output blocks 32 wave + 16 ob + (0..15): in is an LDS activation
// array [32][pitch] (k >= k_valid read as 0), W a BF16 row-major [128][32 KB]
// copy. The 16x16x32 tiles: A = in[16 sh + lane%16][32 kb + 8 (lane/16) + j],
// B = W[n][32 kb + 8 (lane/16) + j] with n = 32 wave + 16 ob + lane%16; the
// result tile acc[sh][ob] holds rows 16 sh + 4 (lane/16) + i of column n.
template <int KB>
__device__ inline void layer_product(const uint16_t* __restrict__ w, const float* in, int pitch, int k_valid,
                                     floatx4 (&acc)[2][2])
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int sh = 0; sh < 2; sh++)
#pragma unroll
        for (int ob = 0; ob < 2; ob++) acc[sh][ob] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < KB; kb++) {
        bf16x8 ah[2], am[2], al[2];
#pragma unroll
        for (int sh = 0; sh < 2; sh++) {
            float v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int k = 32 * kb + 8 * kg + j;
                v[j] = k < k_valid ? in[(16 * sh + m) * pitch + k] : 0.0f;
            }
            split8x3(v, ah[sh], am[sh], al[sh]);
        }
#pragma unroll
        for (int ob = 0; ob < 2; ob++) {
            const int n = 32 * wave + 16 * ob + m;
            const bf16x8 bw = load_w8(w + n * (32 * KB) + 32 * kb + 8 * kg);
#pragma unroll
            for (int sh = 0; sh < 2; sh++)
                acc[sh][ob] = mfma(al[sh], bw, mfma(am[sh], bw, mfma(ah[sh], bw, acc[sh][ob])));
        }
    }
}

// Hidden layer: z[s][o] = b[o] + sum_i W[o][i] in[s][i], h = gelu(z).
template <int KB>
__device__ inline void hidden_layer(const uint16_t* __restrict__ w, const float* __restrict__ b, const float* in,
                                    int pitch, int k_valid, float (*z)[kMlpHidden], float (*h)[kMlpHidden])
{
    floatx4 acc[2][2];
    layer_product<KB>(w, in, pitch, k_valid, acc);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int ob = 0; ob < 2; ob++) {
        const int o = 32 * wave + 16 * ob + m;
        const float bias = b[o];
#pragma unroll
        for (int sh = 0; sh < 2; sh++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int r = 16 * sh + 4 * kg + i;
                const float zz = acc[sh][ob][i] + bias;
                if (z) z[r][o] = zz;
                h[r][o] = gelu(zz);
            }
    }
}

__device__ inline void forward(const MlpDevice& m, Lds& s)
{
    hidden_layer<kMlpInputsPadded / 32>(m.w1, m.master + kMlpOffB1, &s.x[0][0], kXPad, kMlpInputs, s.z1, s.h1);
    __syncthreads();
    hidden_layer<kMlpHidden / 32>(m.w2, m.master + kMlpOffB2, &s.h1[0][0], kMlpHidden, kMlpHidden, s.z2, s.h2);
    __syncthreads();
    if (threadIdx.x < kS * kMlpOutputs) {
        const int r = threadIdx.x >> 1, j = threadIdx.x & 1;
        float acc = m.master[kMlpOffB3 + j];
        for (int i = 0; i < kMlpHidden; i++) acc = fmaf(bf16f(m.w3t[i * kMlpOutputs + j]), s.h2[r][i], acc);
        s.out[r][j] = acc;
    }
    __syncthreads();
}

// Weight gradient dW[o][i] = sum_s d[s][o] a[s][i] over the 32 samples for
// the wave's two 16-row output blocks (o = 32 wave + 16 ob + ...) and the
// input blocks ib < NB (i = 16 ib + ...; i >= i_valid not written): A =
// d[8 (lane/16) + j][o], B = a[8 (lane/16) + j][i], both split, three MFMAs
// per tile. g: the gradient block's row-major [128][i_valid] slot.
template <int NB>
__device__ inline void weight_gradient(const float (*d)[kMlpHidden], const float* a, int pitch, int i_valid,
                                       float* __restrict__ g)
{
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, kg = lane >> 4;
    bf16x8 dh[2], dl[2];
#pragma unroll
    for (int ob = 0; ob < 2; ob++) {
        const int o = 32 * wave + 16 * ob + m;
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = d[8 * kg + j][o];
        split8(v, dh[ob], dl[ob]);
    }
    for (int ib = 0; ib < NB; ib++) {
        const int i = 16 * ib + m;
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = i < i_valid ? a[(8 * kg + j) * pitch + i] : 0.0f;
        bf16x8 ah, al;
        split8(v, ah, al);
#pragma unroll
        for (int ob = 0; ob < 2; ob++) {
            floatx4 c = {0.0f, 0.0f, 0.0f, 0.0f};
            c = mfma(dh[ob], ah, c);
            c = mfma(dl[ob], ah, c);
            c = mfma(dh[ob], al, c);
            if (i < i_valid) {
#pragma unroll
                for (int r = 0; r < 4; r++) g[(32 * wave + 16 * ob + 4 * kg + r) * i_valid + i] = c[r];
            }
        }
    }
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
    // dW2[o][i] = sum_s dz2[s][o] h1[s][i]; db2
    weight_gradient<kMlpHidden / 16>(s.z2, &s.h1[0][0], kMlpHidden, kMlpHidden, g + kMlpOffW2);
    if (t < kMlpHidden) {
        float b = 0.0f;
        for (int r = 0; r < kS; r++) b += s.z2[r][t];
        g[kMlpOffB2 + t] = b;
    }
    {
        // dz1[s][i] = (sum_o dz2[s][o] W2[o][i]) * gelu'(z1[s][i]), into h2: the
        // layer product with W2^T (row i of W2^T is column i of W2)
        floatx4 acc[2][2];
        layer_product<kMlpHidden / 32>(m.w2t, &s.z2[0][0], kMlpHidden, kMlpHidden, acc);
        const int wave = t >> 6, lane = t & 63;
        const int mm = lane & 15, kg = lane >> 4;
#pragma unroll
        for (int ob = 0; ob < 2; ob++) {
            const int i = 32 * wave + 16 * ob + mm;
#pragma unroll
            for (int sh = 0; sh < 2; sh++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int r = 16 * sh + 4 * kg + q;
                    s.h2[r][i] = acc[sh][ob][q] * gelu_grad(s.z1[r][i]);
                }
        }
    }
    __syncthreads();
    // dW1[o][i] = sum_s dz1[s][o] x[s][i] (inputs 55..63 of the last block not written); db1
    weight_gradient<kMlpInputsPadded / 16>(s.h2, &s.x[0][0], kXPad, kMlpInputs, g + kMlpOffW1);
    if (t < kMlpHidden) {
        float b = 0.0f;
        for (int r = 0; r < kS; r++) b += s.h2[r][t];
        g[kMlpOffB1 + t] = b;
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
        m.w1[o * kMlpInputsPadded + i] = h;
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
