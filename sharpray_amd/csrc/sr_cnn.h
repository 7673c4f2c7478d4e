// Kernel-predicting U-Net denoiser (no reference counterpart; the reference denoises with the
// joint bilateral Compositor.Filter, compose.cs:101-204, which stays available beside it).
//
// Input per pixel: log(1 + radiance) (3), albedo (3), normal (3), 1 / (1 + depth) (0 where no
// surface) - 10 channels. Five levels with 32 / 48 / 64 / 96 / 128 channels: two 3x3 conv + ReLU
// per level, 2x2 average pooling down, nearest upsampling and skip concatenation up, and a last
// 3x3 conv to 4 x 25 logits: four predicted 5x5 kernels (softmax-normalised) applied in cascade
// to the noisy radiance with their taps 1, 2, 4 and 8 pixels apart (a 61-pixel footprint).
// Convolutions run as implicit GEMMs on bf16 MFMA (v_mfma_f32_16x16x32_bf16) with fp32
// accumulation; activations are bf16, NHWC with the channel stride padded to a multiple of 32.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include <vector>

namespace sr {

struct CnnLayerSpec {
    int cinA, cinB;  // true input channels (cinB: skip channels of a concatenation)
    int cout;
    int mode;        // 0 same resolution, 1 2x2 average pool of the source, 2 concat(up(source), skip)
    int srcA, srcB;  // source activations: -1 = input features, else a layer index
    int level;       // resolution level of the output (0 = full)
    bool relu;
};

// Stages of the predicted-kernel cascade (taps 2^stage pixels apart).
constexpr int kKpnStages = 4;
// read_layer index of layer L's masked output gradient from the last training step.
constexpr int kGradLayerBase = 1000;
// The fixed topology (19 conv layers).
const std::vector<CnnLayerSpec>& cnn_layers();
// Floats in a weight blob: per layer, weights [cout][cinA + cinB][3][3] then bias [cout].
uint64_t cnn_weight_count();

class CnnNet {
public:
    ~CnnNet();
    // Converts a float weight blob to the padded bf16 device layout. Returns false on a HIP error.
    bool load(const float* weights, hipStream_t s);
    bool loaded() const { return loaded_; }
    // Denoises W x H: rad/alb/nrm (3 floats per pixel), dep (1 float, +inf = no surface), out rgb.
    // All pointers are device memory. Returns false on a HIP error.
    bool run(const float* rad, const float* alb, const float* nrm, const float* dep, uint32_t W, uint32_t H,
             float* out, hipStream_t s);
    // Activations of `layer` from the last run (layer -1: input features) as float, true channels
    // only: out[(y * w + x) * c + k]. Returns false if unavailable.
    bool read_layer(int layer, std::vector<float>& out, uint32_t& c, uint32_t& w, uint32_t& h) const;
    // Algorithmic floating-point operations of one run at W x H (2 per multiply-add, true channels).
    static double flops(uint32_t W, uint32_t H);

    // ---- Training: Adam on fp32 master weights, bf16 copies refreshed after every step. ----------
    // Starts from the loaded weights with zero Adam moments.
    bool train_init(hipStream_t s);
    bool training() const { return !mw_.empty(); }
    // One step on one example (device arrays; target: w*h*3 radiance): forward, loss
    // mean (out - target)^2 / (target^2 + 0.01) over pixels and channels (added to *lossDev),
    // backward, Adam with learning rate lr.
    bool train_step(const float* rad, const float* alb, const float* nrm, const float* dep, const float* target,
                    uint32_t W, uint32_t H, float lr, float* lossDev, hipStream_t s, bool apply = true);
    // Gradients left by a train_step with apply = false, in the blob layout; cleared afterwards.
    bool export_grads(std::vector<float>& blob, hipStream_t s);
    // Master weights in the blob layout of cnn_weight_count().
    bool export_weights(std::vector<float>& blob) const;
    bool export_blob(const std::vector<float*>& w, const std::vector<float*>& b, std::vector<float>& blob) const;

private:
    bool ensure(uint32_t W, uint32_t H);
    bool ensure_train(uint32_t W, uint32_t H);
    void release_acts();
    void release_train_acts();
    bool pack(hipStream_t s);
    // Training state (per layer): fp32 master weights [9][coutPad][cinPad] and biases, their
    // gradients and Adam moments, bf16 flipped-transposed weights [9][cinPad][coutPad] for the
    // data gradient, fp32 activation gradients, bf16 masked output gradients.
    std::vector<float*> mw_, mb_, gw_, gb_, m1w_, m2w_, m1b_, m2b_;
    std::vector<uint16_t*> wt_;
    std::vector<float*> dact_;
    std::vector<uint16_t*> dyp_;
    uint16_t* dx_ = nullptr;
    float* zeros_ = nullptr;
    float* out_ = nullptr;
    uint32_t tW_ = 0, tH_ = 0;
    int step_ = 0;
    bool loaded_ = false;
    std::vector<uint16_t*> w_;  // per layer [9][coutPad][cinPad] bf16
    std::vector<float*> b_;     // per layer [coutPad]
    uint16_t* feat_ = nullptr;  // input features, stride 32
    std::vector<uint16_t*> act_;
    float* casc_[kKpnStages - 1] = {};  // radiance after each cascade stage but the last
    float* gcas_[kKpnStages] = {};      // loss gradients of the cascade stages' outputs
    uint32_t W_ = 0, H_ = 0;
};

}  // namespace sr
