// surrogate.cpp - host side of the score surrogate network (mlp.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "errors.hpp"
#include "mlp.hpp"

namespace vship {

namespace {

void dmalloc(void** p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) VSHIP_THROW(Vship_OutOfVRAM, "hipMalloc failed in the surrogate");
}

}  // namespace

Surrogate::Surrogate(int max_batch, int device) : device_(device), max_batch_(max_batch)
{
    if (max_batch < 1 || max_batch > (1 << 20)) VSHIP_THROW(Vship_BadDeviceArgument, "surrogate max_batch out of range");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    VSHIP_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    dmalloc((void**)&dev_.master, kMlpParams * sizeof(float));
    dmalloc((void**)&dev_.adam_m, kMlpParams * sizeof(float));
    dmalloc((void**)&dev_.adam_v, kMlpParams * sizeof(float));
    dmalloc((void**)&dev_.w1, kMlpHidden * kMlpInputsPadded * sizeof(uint16_t));
    // the padding inputs 55..63 of the BF16 W1 copy stay zero (requantisation writes the rest)
    VSHIP_HIP_CHECK(hipMemset(dev_.w1, 0, kMlpHidden * kMlpInputsPadded * sizeof(uint16_t)));
    dmalloc((void**)&dev_.w2t, kMlpHidden * kMlpHidden * sizeof(uint16_t));
    dmalloc((void**)&dev_.w3t, kMlpOutputs * kMlpHidden * sizeof(uint16_t));
    dmalloc((void**)&dev_.w2, kMlpHidden * kMlpHidden * sizeof(uint16_t));
    dmalloc((void**)&dev_.w3, kMlpOutputs * kMlpHidden * sizeof(uint16_t));
    dmalloc((void**)&d_norm_, 2 * kMlpInputs * sizeof(float));
    const int groups = mlp_groups(max_batch);
    dmalloc((void**)&d_partial_, (size_t)groups * kMlpParams * sizeof(float));
    dmalloc((void**)&d_loss_part_, (size_t)groups * sizeof(float));
    dmalloc((void**)&d_grad_, kMlpParams * sizeof(float));
    dmalloc((void**)&d_loss_, sizeof(float));
    loss_cap_ = 1;
    dev_.xmean = d_norm_;
    dev_.xinv = d_norm_ + kMlpInputs;
    // identity normalisation and zero weights until the caller sets them
    std::vector<float> norm(2 * kMlpInputs, 0.0f);
    for (int i = 0; i < kMlpInputs; i++) norm[kMlpInputs + i] = 1.0f;
    VSHIP_HIP_CHECK(hipMemcpy(d_norm_, norm.data(), norm.size() * sizeof(float), hipMemcpyHostToDevice));
    dev_.ymean = 0.0f;
    dev_.ystd = dev_.yinv = 1.0f;
    std::vector<float> zeros(kMlpParams, 0.0f);
    set_params(zeros.data(), kMlpParams);
}

Surrogate::~Surrogate()
{
    // destructors report nothing: errors here leave nothing to recover
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    void* ptrs[] = {dev_.master, dev_.adam_m, dev_.adam_v, dev_.w1, dev_.w2t, dev_.w3t, dev_.w2, dev_.w3,
                    d_norm_, d_partial_, d_loss_part_, d_grad_, d_loss_, d_idx_};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void Surrogate::set_params(const float* host, int count)
{
    if (count != kMlpParams) VSHIP_THROW(Vship_BadDeviceArgument, "surrogate parameter count differs from the network");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    VSHIP_HIP_CHECK(hipMemcpyAsync(dev_.master, host, kMlpParams * sizeof(float), hipMemcpyHostToDevice, stream_));
    VSHIP_HIP_CHECK(hipMemsetAsync(dev_.adam_m, 0, kMlpParams * sizeof(float), stream_));
    VSHIP_HIP_CHECK(hipMemsetAsync(dev_.adam_v, 0, kMlpParams * sizeof(float), stream_));
    VSHIP_HIP_CHECK(launch_mlp_requant(dev_, stream_));
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
    steps_ = 0;
}

void Surrogate::get_params(float* host, int count)
{
    if (count != kMlpParams) VSHIP_THROW(Vship_BadDeviceArgument, "surrogate parameter count differs from the network");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    VSHIP_HIP_CHECK(hipMemcpyAsync(host, dev_.master, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, stream_));
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Surrogate::set_norm(const float* xmean, const float* xstd, float ymean, float ystd)
{
    if (!(ystd > 0.0f)) VSHIP_THROW(Vship_BadDeviceArgument, "surrogate target standard deviation must be positive");
    std::vector<float> norm(2 * kMlpInputs);
    for (int i = 0; i < kMlpInputs; i++) {
        norm[i] = xmean[i];
        norm[kMlpInputs + i] = xstd[i] > 0.0f ? 1.0f / xstd[i] : 0.0f;
    }
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    VSHIP_HIP_CHECK(hipMemcpyAsync(d_norm_, norm.data(), norm.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
    dev_.ymean = ymean;
    dev_.ystd = ystd;
    dev_.yinv = 1.0f / ystd;
}

void Surrogate::predict(const float* dX, int n, float* dmean, float* dsigma)
{
    if (n < 0) VSHIP_THROW(Vship_BadDeviceArgument, "negative sample count");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    VSHIP_HIP_CHECK(launch_mlp_forward(dev_, dX, n, dmean, dsigma, stream_));
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Surrogate::gradients(const float* dX, const float* dY, int n, float* host_grad, double* loss)
{
    if (n < 1 || n > max_batch_) VSHIP_THROW(Vship_BadDeviceArgument, "gradient batch outside 1..max_batch");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    const int groups = mlp_groups(n);
    VSHIP_HIP_CHECK(launch_mlp_grad(dev_, dX, dY, nullptr, n, 1.0f / (float)n, d_partial_, d_loss_part_, stream_));
    MlpAdam none{};
    VSHIP_HIP_CHECK(launch_mlp_update(dev_, d_partial_, groups, none, d_grad_, d_loss_part_, d_loss_, stream_));
    float l = 0.0f;
    VSHIP_HIP_CHECK(hipMemcpyAsync(host_grad, d_grad_, kMlpParams * sizeof(float), hipMemcpyDeviceToHost, stream_));
    VSHIP_HIP_CHECK(hipMemcpyAsync(&l, d_loss_, sizeof(float), hipMemcpyDeviceToHost, stream_));
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
    if (loss) *loss = (double)l / n;
}

void Surrogate::train(const float* dX, const float* dY, int n, int epochs, int batch, float lr, float weight_decay,
                      uint64_t seed, float* loss_per_epoch)
{
    if (n < 1 || epochs < 1 || batch < 1 || batch > max_batch_ || !(lr > 0.0f))
        VSHIP_THROW(Vship_BadDeviceArgument, "surrogate training arguments out of range");
    VSHIP_HIP_CHECK(hipSetDevice(device_));
    const int steps_per_epoch = (n + batch - 1) / batch;
    const int total_steps = steps_per_epoch * epochs;
    if (loss_cap_ < total_steps) {
        if (d_loss_) (void)hipFree(d_loss_);
        d_loss_ = nullptr;
        dmalloc((void**)&d_loss_, (size_t)total_steps * sizeof(float));
        loss_cap_ = total_steps;
    }
    if (idx_cap_ < n) {
        if (d_idx_) (void)hipFree(d_idx_);
        d_idx_ = nullptr;
        dmalloc((void**)&d_idx_, (size_t)n * sizeof(int));
        idx_cap_ = n;
    }
    std::vector<int> perm(n);
    std::mt19937_64 rng(seed);
    const float beta1 = 0.9f, beta2 = 0.999f;
    int step = 0;
    for (int e = 0; e < epochs; e++) {
        // the previous epoch's copy of perm must have finished before perm
        // is rewritten
        VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
        std::iota(perm.begin(), perm.end(), 0);
        std::shuffle(perm.begin(), perm.end(), rng);
        VSHIP_HIP_CHECK(hipMemcpyAsync(d_idx_, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, stream_));
        for (int s = 0; s < steps_per_epoch; s++, step++) {
            const int base = s * batch;
            const int count = std::min(batch, n - base);
            steps_++;
            MlpAdam a;
            const double progress = total_steps > 1 ? (double)step / (double)(total_steps - 1) : 1.0;
            a.lr = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * progress))));
            a.beta1 = beta1;
            a.beta2 = beta2;
            a.eps = 1e-8f;
            a.weight_decay = weight_decay;
            a.corr1 = (float)(1.0 / (1.0 - std::pow((double)beta1, (double)steps_)));
            a.corr2 = (float)(1.0 / (1.0 - std::pow((double)beta2, (double)steps_)));
            VSHIP_HIP_CHECK(launch_mlp_grad(dev_, dX, dY, d_idx_ + base, count, 1.0f / (float)count, d_partial_,
                                            d_loss_part_, stream_));
            VSHIP_HIP_CHECK(launch_mlp_update(dev_, d_partial_, mlp_groups(count), a, nullptr, d_loss_part_,
                                              d_loss_ + step, stream_));
        }
    }
    std::vector<float> losses(total_steps);
    VSHIP_HIP_CHECK(hipMemcpyAsync(losses.data(), d_loss_, (size_t)total_steps * sizeof(float), hipMemcpyDeviceToHost,
                                   stream_));
    VSHIP_HIP_CHECK(hipStreamSynchronize(stream_));
    if (loss_per_epoch) {
        for (int e = 0; e < epochs; e++) {
            double sum = 0.0;
            for (int s = 0; s < steps_per_epoch; s++) sum += losses[(size_t)e * steps_per_epoch + s];
            loss_per_epoch[e] = (float)(sum / n);
        }
    }
}

}  // namespace vship
