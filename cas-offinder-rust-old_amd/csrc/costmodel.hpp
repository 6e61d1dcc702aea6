// Cost model of the bulge scheduler: chunk features + FP32 MLP (host side of
// csrc/costmodel.hip).  No reference counterpart.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cof_internal.hpp"

namespace cof {

constexpr int kMlpIn = 12;
constexpr int kMlpHidden = 64;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden + kMlpHidden + 1;
static_assert(kMlpParams == 5057, "12-64-64-1");

// Growable device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    void ensure(size_t bytes);
    ~DevBuf();
};

// Query-dependent feature inputs (distinct PAMs, guide 8-mer bitmaps).
struct FeaturePlan;
FeaturePlan* feature_plan_create(const Query& q);
void feature_plan_destroy(FeaturePlan* fp);
// d_out[c * 12 + i] = feature i of genome chunk c
void chunk_features(const FeaturePlan* fp, const Genome& g, const uint32_t* d_genome, hipStream_t st,
                    DevBuf& chunks_buf, float* d_out);

void mlp_forward_dev(const float* d_params, const float* d_x, uint32_t n, float* d_y, hipStream_t st);

// FP32 MLP 12 -> 64 -> 64 -> 1 with ReLU, resident on one device, Adam state.
struct Mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* d_params = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    float* d_grad = nullptr;       // kMlpParams gradient + loss
    uint32_t* d_step = nullptr;
    float* d_part = nullptr;       // per-workgroup partial gradients
    hipGraphExec_t train_exec = nullptr;
    Mlp(int device, uint64_t seed);
    ~Mlp();
    void set_params(const float* p);        // also resets the optimizer
    void get_params(float* p) const;
    void reset_optimizer();
    void forward(const float* x, uint64_t n, float* y);
    void loss_grad(const float* x, const float* t, uint64_t n, float* loss, float* grad);
    void train(const float* x, const float* t, uint64_t n, uint32_t steps, float lr, bool eager, float* loss);
    void launch_grad(const float* d_x, const float* d_t, uint32_t n);
};

}  // namespace cof
