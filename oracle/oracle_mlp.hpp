// CPU oracle of the learned inlier-probability network -- TEST
// INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg as
// the checker of the device kernels; the engine never links it.
//
// What it restates.  The reference has no learned part; its hook is the
// PROSAC sampler, which draws from a prefix of the points ordered by a
// quality score (inc/samplers/prosac_sampler.h:51-237, growth function
// :129-174, sample :207-236), and the drivers' per-point `probabilities`
// argument (py/gcransac_python.cpp:4344-4359, py/bindings.cpp:1003-1022).
// This file defines the quality score that feeds that order:
//   * knn_features: per correspondence, 8 floats
//       [x1 y1 x2 y2 scaled to [-1, 1] by the image sizes,
//        descriptor-ratio score,
//        mean displacement (x2 - x1) of the 8 nearest neighbours in image 1
//        minus the point's own displacement, times 2 / max(w1, h1),
//        0 (padding)]
//     all in float, products and sums in the order written (no FMA);
//     neighbours ordered by (squared distance, index), the point itself
//     excluded;
//   * an MLP 8-128-128-128-1 (input 7 features + a zero pad), GELU (tanh
//     form) on the hidden layers, sigmoid output;  BF16 weights and
//     activations into every product, FP32 accumulation, as the device
//     computes it (round_bf16 marks every rounding point);
//   * binary cross-entropy on logits averaged over the batch, its exact
//     gradient (backward pass) and one Adam step.
//
// The template parameter R selects the arithmetic: RoundBF16 (float, with
// the device's BF16 rounding points; the parity reference) or Exact (double,
// no rounding; the finite-difference check of the backward pass).
//
// Pinned by (tests/test_oracle_mlp.py):
//   * the backward pass against central finite differences of the Exact
//     forward loss on every parameter block;
//   * kNN neighbour sets against scipy's cKDTree and a brute-force sort;
//   * GELU, softplus, sigmoid against closed-form values; the first Adam
//     step against its closed form lr * g / (|g| + eps);
//   * BF16 rounding against numpy's float32 bit manipulation (ties to even).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {
namespace mlp {

constexpr int IN = 8;
constexpr int HID = 128;
constexpr int OFF_W1 = 0;                   // [HID][IN]
constexpr int OFF_B1 = OFF_W1 + HID * IN;   // [HID]
constexpr int OFF_W2 = OFF_B1 + HID;        // [HID][HID]
constexpr int OFF_B2 = OFF_W2 + HID * HID;  // [HID]
constexpr int OFF_W3 = OFF_B2 + HID;        // [HID][HID]
constexpr int OFF_B3 = OFF_W3 + HID * HID;  // [HID]
constexpr int OFF_W4 = OFF_B3 + HID;        // [HID]
constexpr int OFF_B4 = OFF_W4 + HID;        // scalar
constexpr int NPARAM = OFF_B4 + 1;          // 34,305

inline float round_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

struct RoundBF16 {
  using T = float;
  static float q(float v) { return round_bf16(v); }
  static float tanh_(float v) { return std::tanh(v); }
  static float exp_(float v) { return std::exp(v); }
  static float log1p_(float v) { return std::log1p(v); }
};
struct Exact {
  using T = double;
  static double q(double v) { return v; }
  static double tanh_(double v) { return std::tanh(v); }
  static double exp_(double v) { return std::exp(v); }
  static double log1p_(double v) { return std::log1p(v); }
};

// GELU, tanh form: 0.5 z (1 + tanh(c (z + 0.044715 z^3))), c = sqrt(2/pi).
template <class R>
typename R::T gelu(typename R::T z) {
  using T = typename R::T;
  const T c = (T)0.7978845608028654;
  const T u = c * (z + (T)0.044715 * (z * z * z));
  return (T)0.5 * z * ((T)1 + R::tanh_(u));
}
template <class R>
typename R::T gelu_grad(typename R::T z) {
  using T = typename R::T;
  const T c = (T)0.7978845608028654;
  const T z2 = z * z;
  const T t = R::tanh_(c * (z + (T)0.044715 * (z2 * z)));
  return (T)0.5 * ((T)1 + t) + (T)0.5 * z * (((T)1 - t * t) * (c * ((T)1 + (T)0.134145 * z2)));
}
template <class R>
typename R::T softplus(typename R::T z) {
  using T = typename R::T;
  const T a = z < (T)0 ? -z : z;
  return (z > (T)0 ? z : (T)0) + R::log1p_(R::exp_(-a));
}
template <class R>
typename R::T sigmoid(typename R::T z) {
  using T = typename R::T;
  return (T)1 / ((T)1 + R::exp_(-z));
}

// ---- features ------------------------------------------------------------

// corr: n x 4 (x1 y1 x2 y2); ratio: n; feat: n x IN.
inline void knn_features(const double* corr, int64_t n, int w1, int h1, int w2, int h2, const double* ratio,
                         float* feat) {
  std::vector<float> x1((size_t)n), y1((size_t)n), ux((size_t)n), uy((size_t)n);
  const float s1x = 2.0f / (float)w1, s1y = 2.0f / (float)h1, s2x = 2.0f / (float)w2, s2y = 2.0f / (float)h2;
  const float sd = 2.0f / (float)(w1 > h1 ? w1 : h1);
  for (int64_t i = 0; i < n; ++i) {
    x1[i] = (float)corr[4 * i + 0];
    y1[i] = (float)corr[4 * i + 1];
    const float x2 = (float)corr[4 * i + 2], y2 = (float)corr[4 * i + 3];
    ux[i] = x2 - x1[i];
    uy[i] = y2 - y1[i];
    float* f = feat + IN * i;
    f[0] = x1[i] * s1x - 1.0f;
    f[1] = y1[i] * s1y - 1.0f;
    f[2] = x2 * s2x - 1.0f;
    f[3] = y2 * s2y - 1.0f;
    f[4] = (float)ratio[i];
  }
  for (int64_t q = 0; q < n; ++q) {
    float td[8];
    int64_t ti[8];
    for (int k = 0; k < 8; ++k) {
      td[k] = INFINITY;
      ti[k] = -1;
    }
    for (int64_t j = 0; j < n; ++j) {
      if (j == q) continue;
      const float dx = x1[q] - x1[j], dy = y1[q] - y1[j];
      const float d = dx * dx + dy * dy;
      // stable insertion: candidates arrive in ascending index order
      for (int k = 7; k > 0; --k) {
        if (d < td[k - 1]) {
          td[k] = td[k - 1];
          ti[k] = ti[k - 1];
        } else if (d < td[k]) {
          td[k] = d;
          ti[k] = j;
        }
      }
      if (d < td[0]) {
        td[0] = d;
        ti[0] = j;
      }
    }
    float sx = 0.0f, sy = 0.0f;
    int c = 0;
    for (int k = 0; k < 8; ++k) {
      if (ti[k] < 0) break;
      sx = sx + ux[ti[k]];
      sy = sy + uy[ti[k]];
      ++c;
    }
    float* f = feat + IN * q;
    if (c > 0) {
      const float fc = (float)c;
      f[5] = (sx / fc - ux[q]) * sd;
      f[6] = (sy / fc - uy[q]) * sd;
    } else {
      f[5] = 0.0f;
      f[6] = 0.0f;
    }
    f[7] = 0.0f;
  }
}

// k nearest neighbour indices (for the pinning test), n x 8, -1 padded.
inline void knn_indices(const double* corr, int64_t n, int32_t* out) {
  for (int64_t q = 0; q < n; ++q) {
    float td[8];
    int64_t ti[8];
    for (int k = 0; k < 8; ++k) {
      td[k] = INFINITY;
      ti[k] = -1;
    }
    const float xq = (float)corr[4 * q], yq = (float)corr[4 * q + 1];
    for (int64_t j = 0; j < n; ++j) {
      if (j == q) continue;
      const float dx = xq - (float)corr[4 * j], dy = yq - (float)corr[4 * j + 1];
      const float d = dx * dx + dy * dy;
      for (int k = 7; k > 0; --k) {
        if (d < td[k - 1]) {
          td[k] = td[k - 1];
          ti[k] = ti[k - 1];
        } else if (d < td[k]) {
          td[k] = d;
          ti[k] = j;
        }
      }
      if (d < td[0]) {
        td[0] = d;
        ti[0] = j;
      }
    }
    for (int k = 0; k < 8; ++k) out[8 * q + k] = (int32_t)ti[k];
  }
}

// ---- network -------------------------------------------------------------

template <class R>
struct Net {
  using T = typename R::T;
  std::vector<T> p;  // parameters as used (BF16-rounded weights for RoundBF16)

  explicit Net(const float* params) : p(NPARAM) {
    for (int i = 0; i < NPARAM; ++i) {
      const bool weight = (i < OFF_B1) || (i >= OFF_W2 && i < OFF_B2) || (i >= OFF_W3 && i < OFF_B3) ||
                          (i >= OFF_W4 && i < OFF_B4);
      p[i] = weight ? R::q((T)params[i]) : (T)params[i];
    }
  }
  explicit Net(const double* params) : p(NPARAM) {
    for (int i = 0; i < NPARAM; ++i) p[i] = (T)params[i];
  }

  // one layer z = W a + b with W [HID][nin]: products summed in double,
  // rounded once (the device sums them in FP32 inside the matrix unit)
  void layer(const T* W, const T* b, const T* a, int nin, T* z) const {
    for (int i = 0; i < HID; ++i) {
      double s = 0.0;
      for (int k = 0; k < nin; ++k) s += (double)W[i * nin + k] * (double)a[k];
      z[i] = (T)s + b[i];
    }
  }

  struct Acts {
    T x[IN], z1[HID], h1[HID], z2[HID], h2[HID], z3[HID], h3[HID], z4;
  };

  void forward(const float* feat, Acts& A) const {
    for (int k = 0; k < IN; ++k) A.x[k] = R::q((T)feat[k]);
    layer(&p[OFF_W1], &p[OFF_B1], A.x, IN, A.z1);
    for (int i = 0; i < HID; ++i) A.h1[i] = R::q(gelu<R>(A.z1[i]));
    layer(&p[OFF_W2], &p[OFF_B2], A.h1, HID, A.z2);
    for (int i = 0; i < HID; ++i) A.h2[i] = R::q(gelu<R>(A.z2[i]));
    layer(&p[OFF_W3], &p[OFF_B3], A.h2, HID, A.z3);
    for (int i = 0; i < HID; ++i) A.h3[i] = R::q(gelu<R>(A.z3[i]));
    // output unit: two partial sums over the neurons i with (i / 4) even and
    // odd, each in ascending order, then added (the device's lane split)
    T s0 = 0, s1 = 0;
    for (int i = 0; i < HID; ++i) {
      if (((i >> 2) & 1) == 0)
        s0 = s0 + p[OFF_W4 + i] * A.h3[i];
      else
        s1 = s1 + p[OFF_W4 + i] * A.h3[i];
    }
    A.z4 = (s0 + s1) + p[OFF_B4];
  }

  T logit(const float* feat) const {
    Acts A;
    forward(feat, A);
    return A.z4;
  }

  // Mean BCE over the n points; grad (NPARAM, accumulated in double) when
  // non-null.  Returns the loss.
  double loss_grad(const float* feat, const float* labels, int64_t n, double* grad) const {
    if (grad) std::fill(grad, grad + NPARAM, 0.0);
    double loss = 0.0;
    const T inv = (T)1 / (T)n;
    Acts A;
    T d1[HID], d2[HID], d3[HID];
    for (int64_t s = 0; s < n; ++s) {
      forward(feat + IN * s, A);
      const T y = (T)labels[s];
      loss += (double)((softplus<R>(A.z4) - y * A.z4) * inv);
      if (!grad) continue;
      const T d4 = (sigmoid<R>(A.z4) - y) * inv;
      grad[OFF_B4] += (double)d4;
      for (int i = 0; i < HID; ++i) {
        grad[OFF_W4 + i] += (double)d4 * (double)A.h3[i];
        d3[i] = R::q((p[OFF_W4 + i] * d4) * gelu_grad<R>(A.z3[i]));
      }
      for (int i = 0; i < HID; ++i) {
        double t = 0.0;
        for (int k = 0; k < HID; ++k) t += (double)p[OFF_W3 + k * HID + i] * (double)d3[k];
        d2[i] = R::q((T)t * gelu_grad<R>(A.z2[i]));
      }
      for (int i = 0; i < HID; ++i) {
        double t = 0.0;
        for (int k = 0; k < HID; ++k) t += (double)p[OFF_W2 + k * HID + i] * (double)d2[k];
        d1[i] = R::q((T)t * gelu_grad<R>(A.z1[i]));
      }
      for (int k = 0; k < HID; ++k) {
        grad[OFF_B3 + k] += (double)d3[k];
        grad[OFF_B2 + k] += (double)d2[k];
        grad[OFF_B1 + k] += (double)d1[k];
        for (int i = 0; i < HID; ++i) {
          grad[OFF_W3 + k * HID + i] += (double)d3[k] * (double)A.h2[i];
          grad[OFF_W2 + k * HID + i] += (double)d2[k] * (double)A.h1[i];
        }
        for (int c = 0; c < IN; ++c) grad[OFF_W1 + k * IN + c] += (double)d1[k] * (double)A.x[c];
      }
    }
    return loss;
  }
};

// One Adam step (float), t = step number after the update (1 for the first).
inline void adam_step(float* params, float* m, float* v, const float* grad, int64_t t, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
  for (int i = 0; i < NPARAM; ++i) {
    const float g = grad[i];
    m[i] = b1 * m[i] + (1.0f - b1) * g;
    v[i] = b2 * v[i] + (1.0f - b2) * (g * g);
    const float mh = m[i] / c1, vh = v[i] / c2;
    params[i] = params[i] - lr * mh / (std::sqrt(vh) + eps);
  }
}

}  // namespace mlp
}  // namespace orc
