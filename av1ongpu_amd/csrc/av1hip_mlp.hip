/*
 * The superblock cost model: a 24 -> 64 -> 64 -> 1 FP32 perceptron (ReLU)
 * that predicts the dependency queue's work per superblock, in microseconds,
 * from the superblock's syntax features. Training (forward, backward, Adam)
 * and inference are hand-written kernels; the decoder uses the predictions
 * only to order its queue, so pixels do not depend on them.
 *
 * Parameters, 5825 floats: W1[64][24], b1[64], W2[64][64], b2[64], W3[64],
 * b3. Features are standardised with the training set's mean and deviation;
 * the output is the standardised time.
 *
 * No counterpart in the reference: its scheduler orders work by wave only
 * (engine.cpp:1080-1302).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "av1hip.h"

namespace {

constexpr int kIn = AV1H_CM_FEATURES, kH = 64;
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH, kB3 = kW3 + kH;
constexpr int kParams = kB3 + 1;
static_assert(kParams == AV1H_CM_PARAMS, "parameter count");
constexpr int kBatch = 64;   /* samples per step */
constexpr int kThreads = 256;

struct Adam {
  float lr, b1, b2, eps;
  int t0; /* steps taken before this launch */
};

/* One workgroup walks the samples in `perm` order, kBatch at a time:
 * forward and backward with activations in LDS (4 threads per sample, 16
 * neurons each), then every thread owns parameters p = tid + k * 256 and
 * reduces their gradient over the batch. mode 0 adds the gradient of the mean
 * squared error over all n samples to `grad`; mode 1 takes an Adam step per
 * batch (gradient of the batch mean). */
__global__ __launch_bounds__(kThreads) void k_mlp_step(float* params, float* grad, float* am, float* av,
                                                      const float* X, const float* Y, const int32_t* perm, int n,
                                                      int mode, Adam adam, float* loss) {
  __shared__ float w[kParams];
  __shared__ float sx[kBatch][kIn];
  __shared__ float h1[kBatch][kH], h2[kBatch][kH], d1[kBatch][kH], d2[kBatch][kH];
  __shared__ float d3[kBatch], sy[kBatch];
  __shared__ float s_loss;
  const int tid = threadIdx.x;
  for (int p = tid; p < kParams; p += kThreads) w[p] = params[p];
  if (tid == 0) s_loss = 0.f;
  const int s = tid >> 2, q = tid & 3; /* sample, quarter of the neurons */
  int step = adam.t0;
  for (int base = 0; base < n; base += kBatch) {
    const int nb = min(kBatch, n - base);
    __syncthreads();
    for (int i = tid; i < kBatch * kIn; i += kThreads) {
      const int r = i / kIn, c = i - r * kIn;
      sx[r][c] = r < nb ? X[(size_t)perm[base + r] * kIn + c] : 0.f;
    }
    if (tid < kBatch) sy[tid] = tid < nb ? Y[perm[base + tid]] : 0.f;
    __syncthreads();
    for (int o = q * 16; o < q * 16 + 16; ++o) {
      float z = w[kB1 + o];
      for (int i = 0; i < kIn; ++i) z += w[kW1 + o * kIn + i] * sx[s][i];
      h1[s][o] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    for (int o = q * 16; o < q * 16 + 16; ++o) {
      float z = w[kB2 + o];
      for (int i = 0; i < kH; ++i) z += w[kW2 + o * kH + i] * h1[s][i];
      h2[s][o] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    if (q == 0) {
      float out = w[kB3];
      for (int i = 0; i < kH; ++i) out += w[kW3 + i] * h2[s][i];
      const float e = s < nb ? out - sy[s] : 0.f;
      const float scale = mode == 0 ? 2.f / (float)n : 2.f / (float)nb;
      d3[s] = e * scale;
      if (s < nb) atomicAdd(&s_loss, e * e);
    }
    __syncthreads();
    for (int o = q * 16; o < q * 16 + 16; ++o) d2[s][o] = h2[s][o] > 0.f ? w[kW3 + o] * d3[s] : 0.f;
    __syncthreads();
    for (int i = q * 16; i < q * 16 + 16; ++i) {
      float g = 0.f;
      if (h1[s][i] > 0.f)
        for (int o = 0; o < kH; ++o) g += w[kW2 + o * kH + i] * d2[s][o];
      d1[s][i] = g;
    }
    __syncthreads();
    ++step;
    for (int p = tid; p < kParams; p += kThreads) {
      float g = 0.f;
      if (p < kB1) {
        const int o = p / kIn, i = p - o * kIn;
        for (int b = 0; b < nb; ++b) g += d1[b][o] * sx[b][i];
      } else if (p < kW2) {
        for (int b = 0; b < nb; ++b) g += d1[b][p - kB1];
      } else if (p < kB2) {
        const int o = (p - kW2) / kH, i = (p - kW2) - o * kH;
        for (int b = 0; b < nb; ++b) g += d2[b][o] * h1[b][i];
      } else if (p < kW3) {
        for (int b = 0; b < nb; ++b) g += d2[b][p - kB2];
      } else if (p < kB3) {
        for (int b = 0; b < nb; ++b) g += d3[b] * h2[b][p - kW3];
      } else {
        for (int b = 0; b < nb; ++b) g += d3[b];
      }
      if (mode == 0) {
        grad[p] += g;
      } else {
        const float m = adam.b1 * am[p] + (1.f - adam.b1) * g;
        const float v = adam.b2 * av[p] + (1.f - adam.b2) * g * g;
        am[p] = m;
        av[p] = v;
        const float mh = m / (1.f - powf(adam.b1, (float)step));
        const float vh = v / (1.f - powf(adam.b2, (float)step));
        w[p] -= adam.lr * mh / (sqrtf(vh) + adam.eps);
      }
    }
  }
  __syncthreads();
  if (mode == 1)
    for (int p = tid; p < kParams; p += kThreads) params[p] = w[p];
  if (tid == 0) *loss = s_loss / (float)n;
}

/* Inference, one thread per sample: standardise, forward, back to
 * microseconds. */
__global__ __launch_bounds__(kThreads) void k_mlp_infer(const float* params, const float* norm, const float* X, int n,
                                                       float* out) {
  __shared__ float w[kParams];
  __shared__ float nm[2 * kIn + 2];
  for (int p = threadIdx.x; p < kParams; p += kThreads) w[p] = params[p];
  for (int p = threadIdx.x; p < 2 * kIn + 2; p += kThreads) nm[p] = norm[p];
  __syncthreads();
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  float x[kIn];
#pragma unroll
  for (int k = 0; k < kIn; ++k) x[k] = (X[(size_t)i * kIn + k] - nm[k]) / nm[kIn + k];
  float h[kH];
#pragma unroll
  for (int o = 0; o < kH; ++o) {
    float z = w[kB1 + o];
#pragma unroll
    for (int k = 0; k < kIn; ++k) z += w[kW1 + o * kIn + k] * x[k];
    h[o] = z > 0.f ? z : 0.f;
  }
  float y = w[kB3];
  for (int o = 0; o < kH; ++o) {
    float z = w[kB2 + o];
#pragma unroll
    for (int k = 0; k < kH; ++k) z += w[kW2 + o * kH + k] * h[k];
    y += w[kW3 + o] * (z > 0.f ? z : 0.f);
  }
  out[i] = y * nm[2 * kIn + 1] + nm[2 * kIn];
}

}  // namespace

/* The allocation-free prediction's staging (av1h_cm_predict_noalloc): the
 * features of up to AV1H_CM_NOALLOC_MAX superblocks, then their predictions.
 * Each device holds its own copy from library load to process end; host code
 * reaches it only through the address hipGetSymbolAddress gives per device. */
static __device__ __align__(256) float scratch[AV1H_CM_NOALLOC_MAX * (AV1H_CM_FEATURES + 1)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch[kMaxDevices] = {};

/* the scratch array's address on `device`, looked up once per device with
 * that device current (at model creation) */
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return AV1H_ERR_INVALID;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return AV1H_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return AV1H_ERR_DEVICE;
  g_scratch[device] = static_cast<float*>(ptr);
  return AV1H_OK;
}

float* scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}
}  // namespace

struct av1h_costmodel {
  int device = 0;
  hipStream_t stream = nullptr;
  float* params = nullptr; /* device */
  float* adam = nullptr;   /* m then v */
  float* norm = nullptr;   /* mean[24], std[24], y mean, y std */
  int steps = 0;
  std::vector<float> h_norm = std::vector<float>(2 * kIn + 2, 0.f);
};

namespace {

#define CM_OK(call)                         \
  do {                                      \
    if ((call) != hipSuccess) return AV1H_ERR_DEVICE; \
  } while (0)

int upload_floats(av1h_costmodel* m, float** dst, const float* src, size_t n) {
  CM_OK(hipMalloc(dst, n * sizeof(float)));
  CM_OK(hipMemcpy(*dst, src, n * sizeof(float), hipMemcpyHostToDevice));
  return AV1H_OK;
}

}  // namespace

extern "C" {

int av1h_cm_create(int device, uint64_t seed, av1h_costmodel** out) {
  if (!out) return AV1H_ERR_INVALID;
  *out = nullptr;
  av1h_device_info d;
  const int rc = av1h_describe_device(device, &d);
  if (rc != AV1H_OK) return rc;
  av1h_costmodel* m = new av1h_costmodel;
  m->device = device;
  if (hipSetDevice(device) != hipSuccess || scratch_setup(device) != AV1H_OK ||
      hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&m->params, kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->adam, 2 * kParams * sizeof(float)) != hipSuccess ||
      hipMalloc(&m->norm, (2 * kIn + 2) * sizeof(float)) != hipSuccess) {
    av1h_cm_destroy(m);
    return AV1H_ERR_DEVICE;
  }
  /* He initialisation, zero biases */
  std::vector<float> p(kParams, 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> g1(0.f, std::sqrt(2.f / kIn)), g2(0.f, std::sqrt(2.f / kH));
  for (int i = 0; i < kH * kIn; ++i) p[kW1 + i] = g1(rng);
  for (int i = 0; i < kH * kH; ++i) p[kW2 + i] = g2(rng);
  for (int i = 0; i < kH; ++i) p[kW3 + i] = g2(rng);
  for (int k = 0; k < kIn; ++k) m->h_norm[kIn + k] = 1.f;
  m->h_norm[2 * kIn + 1] = 1.f;
  if (av1h_cm_set_params(m, p.data()) != AV1H_OK ||
      av1h_cm_set_norm(m, m->h_norm.data(), m->h_norm.data() + kIn, 0.f, 1.f) != AV1H_OK) {
    av1h_cm_destroy(m);
    return AV1H_ERR_DEVICE;
  }
  *out = m;
  return AV1H_OK;
}

void av1h_cm_destroy(av1h_costmodel* m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (void* p : {(void*)m->params, (void*)m->adam, (void*)m->norm})
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

int av1h_cm_set_params(av1h_costmodel* m, const float* params) {
  if (!m || !params) return AV1H_ERR_INVALID;
  (void)hipSetDevice(m->device);
  CM_OK(hipMemcpy(m->params, params, kParams * sizeof(float), hipMemcpyHostToDevice));
  CM_OK(hipMemset(m->adam, 0, 2 * kParams * sizeof(float)));
  m->steps = 0;
  return AV1H_OK;
}

int av1h_cm_get_params(av1h_costmodel* m, float* params) {
  if (!m || !params) return AV1H_ERR_INVALID;
  (void)hipSetDevice(m->device);
  CM_OK(hipMemcpy(params, m->params, kParams * sizeof(float), hipMemcpyDeviceToHost));
  return AV1H_OK;
}

int av1h_cm_set_norm(av1h_costmodel* m, const float* mean, const float* dev, float y_mean, float y_dev) {
  if (!m || !mean || !dev || !(y_dev > 0.f)) return AV1H_ERR_INVALID;
  for (int k = 0; k < kIn; ++k) {
    if (!(dev[k] > 0.f)) return AV1H_ERR_INVALID;
    m->h_norm[k] = mean[k];
    m->h_norm[kIn + k] = dev[k];
  }
  m->h_norm[2 * kIn] = y_mean;
  m->h_norm[2 * kIn + 1] = y_dev;
  (void)hipSetDevice(m->device);
  CM_OK(hipMemcpy(m->norm, m->h_norm.data(), (2 * kIn + 2) * sizeof(float), hipMemcpyHostToDevice));
  return AV1H_OK;
}

int av1h_cm_get_norm(av1h_costmodel* m, float* mean, float* dev, float* y_mean, float* y_dev) {
  if (!m) return AV1H_ERR_INVALID;
  for (int k = 0; k < kIn; ++k) {
    if (mean) mean[k] = m->h_norm[k];
    if (dev) dev[k] = m->h_norm[kIn + k];
  }
  if (y_mean) *y_mean = m->h_norm[2 * kIn];
  if (y_dev) *y_dev = m->h_norm[2 * kIn + 1];
  return AV1H_OK;
}

int av1h_cm_predict(av1h_costmodel* m, const float* features, int64_t n, float* out_us) {
  if (!m || n < 0 || (n && (!features || !out_us))) return AV1H_ERR_INVALID;
  if (n == 0) return AV1H_OK;
  (void)hipSetDevice(m->device);
  float *dx = nullptr, *dy = nullptr;
  int rc = upload_floats(m, &dx, features, (size_t)n * kIn);
  if (rc == AV1H_OK && hipMalloc(&dy, (size_t)n * sizeof(float)) != hipSuccess) rc = AV1H_ERR_DEVICE;
  if (rc == AV1H_OK) {
    hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, m->stream,
                       m->params, m->norm, dx, (int)n, dy);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out_us, dy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess)
      rc = AV1H_ERR_DEVICE;
  }
  if (dx) (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  return rc;
}

/* av1h_cm_predict without an allocation: the features and predictions are
 * staged in the device's scratch array, set up when the model was created.
 * Calls on one device share the array, so calls that overlap there race. */
int av1h_cm_predict_noalloc(av1h_costmodel* m, const float* features, int64_t n, float* out_us) {
  if (!m || n < 0 || (n && (!features || !out_us))) return AV1H_ERR_INVALID;
  if (n > AV1H_CM_NOALLOC_MAX) return AV1H_ERR_UNSUPPORTED;
  if (n == 0) return AV1H_OK;
  (void)hipSetDevice(m->device);
  float* dx = scratch_of(m->device);
  if (!dx) return AV1H_ERR_DEVICE;
  float* dy = dx + (size_t)AV1H_CM_NOALLOC_MAX * kIn;
  CM_OK(hipMemcpyAsync(dx, features, (size_t)n * kIn * sizeof(float), hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, m->stream,
                     m->params, m->norm, dx, (int)n, dy);
  CM_OK(hipGetLastError());
  CM_OK(hipMemcpyAsync(out_us, dy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  CM_OK(hipStreamSynchronize(m->stream));
  return AV1H_OK;
}

/* Standardised copies of the samples, with the model's normalisation. */
static void standardise(const av1h_costmodel* m, const float* X, const float* y_us, int64_t n, std::vector<float>* xs,
                        std::vector<float>* ys) {
  xs->resize((size_t)n * kIn);
  ys->resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < kIn; ++k)
      (*xs)[(size_t)i * kIn + k] = (X[(size_t)i * kIn + k] - m->h_norm[k]) / m->h_norm[kIn + k];
    (*ys)[(size_t)i] = (y_us[i] - m->h_norm[2 * kIn]) / m->h_norm[2 * kIn + 1];
  }
}

static int run_steps(av1h_costmodel* m, const std::vector<float>& xs, const std::vector<float>& ys,
                     const std::vector<int32_t>& perm, int mode, Adam adam, float* grad_host, float* loss_host) {
  const int64_t n = (int64_t)ys.size();
  float *dx = nullptr, *dy = nullptr, *dg = nullptr, *dl = nullptr;
  int32_t* dp = nullptr;
  int rc = upload_floats(m, &dx, xs.data(), xs.size());
  if (rc == AV1H_OK) rc = upload_floats(m, &dy, ys.data(), ys.size());
  if (rc == AV1H_OK && (hipMalloc(&dp, perm.size() * 4) != hipSuccess ||
                        hipMemcpy(dp, perm.data(), perm.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
                        hipMalloc(&dg, kParams * sizeof(float)) != hipSuccess ||
                        hipMemset(dg, 0, kParams * sizeof(float)) != hipSuccess ||
                        hipMalloc(&dl, sizeof(float)) != hipSuccess))
    rc = AV1H_ERR_DEVICE;
  if (rc == AV1H_OK) {
    hipLaunchKernelGGL(k_mlp_step, dim3(1), dim3(kThreads), 0, m->stream, m->params, dg, m->adam,
                       m->adam + kParams, dx, dy, dp, (int)n, mode, adam, dl);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(m->stream) != hipSuccess) rc = AV1H_ERR_DEVICE;
  }
  if (rc == AV1H_OK && grad_host &&
      hipMemcpy(grad_host, dg, kParams * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    rc = AV1H_ERR_DEVICE;
  if (rc == AV1H_OK && loss_host && hipMemcpy(loss_host, dl, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    rc = AV1H_ERR_DEVICE;
  for (void* p : {(void*)dx, (void*)dy, (void*)dg, (void*)dl, (void*)dp})
    if (p) (void)hipFree(p);
  return rc;
}

int av1h_cm_gradients(av1h_costmodel* m, const float* features, const float* y_us, int64_t n, float* grad,
                      float* loss) {
  if (!m || !features || !y_us || n <= 0 || n > (1 << 30) || !grad) return AV1H_ERR_INVALID;
  (void)hipSetDevice(m->device);
  std::vector<float> xs, ys;
  standardise(m, features, y_us, n, &xs, &ys);
  std::vector<int32_t> perm((size_t)n);
  for (int64_t i = 0; i < n; ++i) perm[(size_t)i] = (int32_t)i;
  Adam a{0.f, 0.f, 0.f, 0.f, 0};
  return run_steps(m, xs, ys, perm, 0, a, grad, loss);
}

int av1h_cm_train(av1h_costmodel* m, const float* features, const float* y_us, int64_t n, int epochs, float lr,
                  uint64_t seed, int set_norm, float* loss_per_epoch) {
  if (!m || !features || !y_us || n <= 1 || n > (1 << 30) || epochs < 1 || !(lr > 0.f)) return AV1H_ERR_INVALID;
  (void)hipSetDevice(m->device);
  if (set_norm) {
    std::vector<double> mu(kIn, 0.0), sd(kIn, 0.0);
    double ym = 0, yd = 0;
    for (int64_t i = 0; i < n; ++i) {
      for (int k = 0; k < kIn; ++k) mu[k] += features[(size_t)i * kIn + k];
      ym += y_us[i];
    }
    for (int k = 0; k < kIn; ++k) mu[k] /= (double)n;
    ym /= (double)n;
    for (int64_t i = 0; i < n; ++i) {
      for (int k = 0; k < kIn; ++k) {
        const double d = features[(size_t)i * kIn + k] - mu[k];
        sd[k] += d * d;
      }
      yd += (y_us[i] - ym) * (y_us[i] - ym);
    }
    std::vector<float> fm(kIn), fs(kIn);
    for (int k = 0; k < kIn; ++k) {
      fm[k] = (float)mu[k];
      fs[k] = (float)std::max(std::sqrt(sd[k] / (double)n), 1e-3);
    }
    const int rc = av1h_cm_set_norm(m, fm.data(), fs.data(), (float)ym, (float)std::max(std::sqrt(yd / (double)n), 1e-3));
    if (rc != AV1H_OK) return rc;
  }
  std::vector<float> xs, ys;
  standardise(m, features, y_us, n, &xs, &ys);
  std::vector<int32_t> perm((size_t)n);
  std::mt19937_64 rng(seed);
  for (int e = 0; e < epochs; ++e) {
    for (int64_t i = 0; i < n; ++i) perm[(size_t)i] = (int32_t)i;
    std::shuffle(perm.begin(), perm.end(), rng);
    Adam a{lr, 0.9f, 0.999f, 1e-8f, m->steps};
    float loss = 0.f;
    const int rc = run_steps(m, xs, ys, perm, 1, a, nullptr, &loss);
    if (rc != AV1H_OK) return rc;
    m->steps += (int)((n + kBatch - 1) / kBatch);
    if (loss_per_epoch) loss_per_epoch[e] = loss;
  }
  return AV1H_OK;
}

}  // extern "C"
