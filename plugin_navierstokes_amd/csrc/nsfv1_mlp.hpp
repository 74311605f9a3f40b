// nsfv1_mlp.hpp -- sizes, parameter layout and launchers of the learned
// initial-guess MLP (24 -> 128 -> 128 -> 4, ReLU, BF16 x BF16 -> FP32).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsfv1 {

constexpr int MLP_IN = 24, MLP_KIN = 32;   // inputs, padded to two MFMA k-blocks
constexpr int MLP_H = 128;                 // hidden width
constexpr int MLP_OUT = 4, MLP_KOUT = 16;  // outputs, padded to the MFMA row block
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_KIN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_KOUT * MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + MLP_KOUT;  // 22800
constexpr int MLP_WG_BLOCKS = 256;                 // workgroups of the weight-gradient reduction
constexpr int PRED_SCALE_BLOCKS = 512;             // workgroups of the per-step scale reduction
constexpr float PRED_CLAMP = 64.f;                 // bound of every scaled feature and target

int64_t mlp_work_floats(int64_t n);
hipError_t launch_mlp_forward(const float* P, const uint16_t* WB, const float* X, int64_t n, float* Y, float* Z1,
                              float* Z2, hipStream_t s);
hipError_t launch_mlp_backward(const uint16_t* WB, const float* X, const float* Z1, const float* Z2, const float* DY,
                               int64_t n, float* grad, float* work, hipStream_t s);
hipError_t launch_mlp_mse(const float* Y, const float* T, int64_t n, float* DY, float* lpart, int* nparts,
                          hipStream_t s);
hipError_t launch_mlp_adam(float* P, float* M, float* V, const float* G, int step, float lr, float b1, float b2,
                           float eps, uint16_t* WB, hipStream_t s);
hipError_t launch_mlp_pack(const float* P, uint16_t* WB, hipStream_t s);
hipError_t launch_predict_scale(const double* u, const double* u1, const double* u2, const double* u3, int64_t nv,
                                double* part, double* sigma, hipStream_t s);
hipError_t launch_predict_features(const int64_t* rowptr, const int32_t* colidx, const double* u, const double* u1,
                                   const double* u2, const double* u3, int64_t nv, const float* scale,
                                   const double* sigma, float* X, hipStream_t s);
hipError_t launch_predict_guess(const double* u, const double* u1, const double* u2, const float* Y,
                                const float* oscale, const double* sigma, int64_t nv, double* g, hipStream_t s);
hipError_t launch_predict_target(const double* un1, const double* u, const double* u1, const double* u2,
                                 const float* oscale, const double* sigma, int64_t nv, float* T, hipStream_t s);

}  // namespace nsfv1
