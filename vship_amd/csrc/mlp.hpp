// mlp.hpp - the score surrogate: a 55 -> 128 -> 128 -> (mean, log-variance)
// GELU network with BF16 weights and FP32 accumulation, trained and run by
// the kernels in mlp.hip (its layer products on the BF16 matrix cores).
//
// Inputs: the 54 SSIMULACRA2 features of three consecutive scales (0-2 or
// 2-4; plane * 18 + scale * 6 + {ssim L1, artifact L1, detail L1, ssim L4,
// artifact L4, detail L4}) and the natural log of the pixel count. Output: the predicted score and its
// standard deviation. Inputs and target are normalised with per-column mean
// and standard deviation fixed before training.
//
// Parameter vector (FP32 master, this order): W1[128][55], b1[128],
// W2[128][128], b2[128], W3[2][128], b3[2] (row = output unit).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>

namespace vship {

constexpr int kMlpInputs = 55;
constexpr int kMlpHidden = 128;
constexpr int kMlpOutputs = 2;
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpHidden * kMlpInputs;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHidden;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHidden;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpOutputs * kMlpHidden;
constexpr int kMlpParams = kMlpOffB3 + kMlpOutputs;  // 23938
constexpr int kMlpSamplesPerGroup = 32;
constexpr int kMlpInputsPadded = 64;  // row pitch of the BF16 W1 copy (inputs 55..63 are zero)
constexpr int kMlpThreads = 256;

// Device state passed by value to the kernels.
struct MlpDevice {
    float* master;     // kMlpParams FP32 weights
    float* adam_m;     // Adam first moment
    float* adam_v;     // Adam second moment
    uint16_t* w1;      // BF16 W1 [128][kMlpInputsPadded], zero past input 55
    uint16_t* w2;      // BF16 W2 [128][128] (forward products)
    uint16_t* w2t;     // BF16 W2^T [128][128] (input gradients of layer 2)
    uint16_t* w3;      // BF16 W3 [2][128]
    uint16_t* w3t;     // BF16 W3^T [128][2]
    const float* xmean;  // kMlpInputs
    const float* xinv;   // 1 / standard deviation (0 for constant columns)
    float ymean, ystd, yinv;
};

struct MlpAdam {
    float lr, beta1, beta2, eps, weight_decay;
    float corr1, corr2;  // 1 / (1 - beta^t)
};

int mlp_groups(int n);
hipError_t launch_mlp_grad(const MlpDevice& m, const float* X, const float* Y, const int* idx, int n, float grad_scale,
                           float* partial, float* loss_partial, hipStream_t stream);
hipError_t launch_mlp_update(const MlpDevice& m, const float* partial, int nparts, const MlpAdam& adam, float* grad,
                             const float* loss_partial, float* loss_out, hipStream_t stream);
hipError_t launch_mlp_requant(const MlpDevice& m, hipStream_t stream);
hipError_t launch_mlp_forward(const MlpDevice& m, const float* X, int n, float* mean, float* sigma,
                              hipStream_t stream);

// Host owner of one network on one device.
class Surrogate {
public:
    Surrogate(int max_batch, int device);
    ~Surrogate();
    Surrogate(const Surrogate&) = delete;
    Surrogate& operator=(const Surrogate&) = delete;

    int device() const { return device_; }
    void set_params(const float* host, int count);      // resets the Adam state
    void get_params(float* host, int count);
    void set_norm(const float* xmean, const float* xstd, float ymean, float ystd);
    // device X [n][55] -> device mean[n], sigma[n] (score units)
    void predict(const float* dX, int n, float* dmean, float* dsigma);
    // summed gradient of the mean loss over n <= max_batch samples (no update)
    void gradients(const float* dX, const float* dY, int n, float* host_grad, double* loss);
    // Adam over `epochs` shuffled passes in steps of `batch` samples; the
    // learning rate decays with a cosine from lr to lr / 20. loss_per_epoch
    // (host, epochs entries) receives the mean training loss of each epoch.
    void train(const float* dX, const float* dY, int n, int epochs, int batch, float lr, float weight_decay,
               uint64_t seed, float* loss_per_epoch);

private:
    int device_;
    int max_batch_;
    hipStream_t stream_ = nullptr;
    MlpDevice dev_{};
    float* d_norm_ = nullptr;     // xmean[55], xinv[55]
    float* d_partial_ = nullptr;  // groups * kMlpParams
    float* d_loss_part_ = nullptr;
    float* d_grad_ = nullptr;
    float* d_loss_ = nullptr;     // per-step loss sums of a train call
    int loss_cap_ = 0;
    int* d_idx_ = nullptr;
    int idx_cap_ = 0;
    int64_t steps_ = 0;           // Adam step count
};

}  // namespace vship
