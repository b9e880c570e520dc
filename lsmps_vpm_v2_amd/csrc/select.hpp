// select.hpp -- learned order / tree-level selection: model layout and the
// host entry points of select.hip.
#pragma once

#include "ctx.hpp"

namespace vpm {

constexpr int kSelFeatures = 9;   // feature row per level (see select.hip)
constexpr int kSelHidden = 64;    // two hidden layers of 64 tanh units
constexpr int kSelParams = kSelFeatures * kSelHidden + kSelHidden + kSelHidden * kSelHidden +
                           kSelHidden + 2 * kSelHidden + 2;
constexpr int kSelRaw = 5;        // raw statistics per level: cells, particles, sum|q|, max|q|, leaves

// Normalisation, training range and selection settings of a model.
struct SelModel {
    float mu[kSelFeatures], sd[kSelFeatures];   // x_norm = (x - mu) / sd
    float lo[kSelFeatures], hi[kSelFeatures];   // training range of each feature
    float margin;      // allowed excursion outside [lo, hi], in units of hi - lo
    float target_err;  // relative L2 velocity error the selection aims at
    int p_min, p_max;  // candidate orders (p_max - p_min < 16, p_max <= 16)
};

// Scratch of one training / gradient evaluation.
struct MlpWork {
    DBuf b_out, b_dout, b_h1, b_h2, b_dh1, b_dh2, b_gpart, b_lpart;
    float *out = nullptr, *dout = nullptr, *h1 = nullptr, *h2 = nullptr, *dh1 = nullptr,
          *dh2 = nullptr, *gpart = nullptr, *lpart = nullptr;
    void reserve(int rows, int nseg);
};

// Level statistics of the built tree in c: raw[(L) * kSelRaw + k] and feature
// rows feats[L * kSelFeatures + i] for L = 0 .. kMaxDepth (order column 0).
void level_features(Ctx &c, int threshold, float *feats_dev, float *raw_dev);
// Orders per level into c.hdr_d->sel_* and orders_dev[0 .. kMaxDepth + 1]
// (one block, after level_features).
void select_orders(Ctx &c, const SelModel *model_dev, const float *params_dev,
                   const float *feats_dev, int fixed_p, uint8_t *orders_dev);

void mlp_forward(hipStream_t s, const float *params, const float *x, int rows, float *out,
                 float *h1, float *h2);
void mlp_loss_grad(hipStream_t s, const float *params, const float *x, int rows, const int *seg,
                   int nseg, const float *y, float w_time, MlpWork &wk, float *grad, float *loss);
void mlp_adam(hipStream_t s, float *params, const float *grad, float *m, float *v, int n,
              int step, float lr, float b1, float b2, float eps);

}  // namespace vpm
