// Table-size predictor (message-passing network): device state shared by
// mpnn.hip and the C ABI.  Not part of the public ABI.
#pragma once
#include "device.hpp"

namespace cuts {

constexpr int kMpnnNodes = 16;   // query vertices per sample (padded)
constexpr int kMpnnFeats = 24;   // input features per vertex
constexpr int kMpnnHidden = 64;  // hidden width = wave64 lanes
constexpr int kMpnnParams = 2 * kMpnnFeats * kMpnnHidden + kMpnnHidden +
                            2 * (2 * kMpnnHidden * kMpnnHidden + kMpnnHidden) + 2 * kMpnnHidden * kMpnnHidden +
                            kMpnnHidden + kMpnnHidden + 1;

struct Mpnn {
  Context* ctx = nullptr;
  float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 master, Adam moments, gradient
  uint16_t *w = nullptr, *wt = nullptr;                            // BF16 copies (input-major, transposed)
  float *part = nullptr, *loss = nullptr;
  uint32_t nblocks = 0;
  uint32_t step = 0;
  // staging for one batch of samples
  uint32_t cap = 0;
  float* x = nullptr;
  uint16_t* adj = nullptr;
  uint32_t* meta = nullptr;
  float *t = nullptr, *y = nullptr;
  ~Mpnn();
};

void mpnn_init(Context& c, Mpnn& n, const float* params);
void mpnn_set(Mpnn& n, const float* params);
void mpnn_get(Mpnn& n, float* params);
void mpnn_forward(Mpnn& n, uint32_t cnt, const float* x, const uint16_t* adj, const uint32_t* meta, float* y);
double mpnn_gradient(Mpnn& n, uint32_t cnt, const float* x, const uint16_t* adj, const uint32_t* meta,
                     const float* t, float* grad_out);
double mpnn_train_step(Mpnn& n, uint32_t cnt, const float* x, const uint16_t* adj, const uint32_t* meta,
                       const float* t, float lr);

}  // namespace cuts
