// pac_mlp.hip -- the learned strategy selector: a 16 -> 64 -> 64 -> 3 ReLU MLP whose
// inference, backward pass and Adam update are hand-written HIP kernels.
//
// No reference counterpart: the reference picks the intersection method from the
// command line (test.cpp:247-259).  The MLP only chooses which kernel runs on a
// partition, so every count stays exact whatever it predicts.
//
// Mapping onto gfx950: one wave per sample with hidden unit j on lane j (64 hidden
// units = one wave64).  Parameters live on the device as bf16 and are widened to
// fp32 in LDS; every dot product accumulates in fp32.  Layer 2 reads the other
// lanes' activations with __shfl, the 3 logits are wave reductions.  The backward
// kernel keeps each lane's parameter rows' gradients in registers across the samples
// its wave handles and writes one partial per wave; a second kernel sums the partials
// in a fixed order (deterministic).  Training keeps an fp32 master copy that Adam
// updates; the bf16 copy used by the forward/backward kernels is re-rounded after
// each step (round to nearest even).
#include <cmath>
#include <cstring>

#include "pac_internal.hpp"

namespace pac {
namespace {

constexpr int kIn = PAC_MLP_IN, kH = PAC_MLP_HIDDEN, kOut = PAC_MLP_OUT;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kOut * kH, kNP = oB3 + kOut;
static_assert(kNP == PAC_MLP_NPARAMS, "parameter layout");

__host__ __device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float wsum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Shared by forward and backward: widen bf16 params into LDS.
__device__ __forceinline__ void load_params(const uint16_t *__restrict__ pb, float *sp) {
  for (int i = threadIdx.x; i < kNP; i += blockDim.x) sp[i] = bf16_to_f32(pb[i]);
  __syncthreads();
}

// Forward for the sample of this wave; lane j returns h1_j, h2_j and all logits.
__device__ __forceinline__ void fwd(const float *sp, const float *__restrict__ x, int lane,
                                    float &h1, float &h2, float z[kOut]) {
  float xi[kIn];
#pragma unroll
  for (int i = 0; i < kIn; ++i) xi[i] = x[i];
  float s = sp[oB1 + lane];
#pragma unroll
  for (int i = 0; i < kIn; ++i) s = fmaf(sp[oW1 + lane * kIn + i], xi[i], s);
  h1 = s > 0.f ? s : 0.f;
  s = sp[oB2 + lane];
#pragma unroll 8
  for (int i = 0; i < kH; ++i) s = fmaf(sp[oW2 + lane * kH + i], __shfl(h1, i, 64), s);
  h2 = s > 0.f ? s : 0.f;
#pragma unroll
  for (int o = 0; o < kOut; ++o) z[o] = wsum(sp[oW3 + o * kH + lane] * h2) + sp[oB3 + o];
}

__global__ __launch_bounds__(256) void k_mlp_forward(const uint16_t *__restrict__ pb,
                                                     const float *__restrict__ x, uint64_t n,
                                                     float *__restrict__ logits,
                                                     uint8_t *__restrict__ choice) {
  __shared__ float sp[kNP];
  load_params(pb, sp);
  const int lane = threadIdx.x & 63;
  const uint64_t wpb = blockDim.x >> 6;
  for (uint64_t s = blockIdx.x * wpb + (threadIdx.x >> 6); s < n; s += gridDim.x * wpb) {
    float h1, h2, z[kOut];
    fwd(sp, x + s * kIn, lane, h1, h2, z);
    if (lane == 0) {
      if (logits)
        for (int o = 0; o < kOut; ++o) logits[s * kOut + o] = z[o];
      if (choice) {
        int best = 0;
        for (int o = 1; o < kOut; ++o)
          if (z[o] > z[best]) best = o;
        choice[s] = (uint8_t)best;
      }
    }
  }
}

// Per-wave partial gradients of the (weighted) mean softmax cross-entropy.
// partial layout: [wave][kNP + 1] with the loss in the last slot.
__global__ __launch_bounds__(256) void k_mlp_backward(const uint16_t *__restrict__ pb,
                                                      const float *__restrict__ x,
                                                      const int32_t *__restrict__ y,
                                                      const float *__restrict__ wts, uint64_t n,
                                                      float *__restrict__ partial) {
  __shared__ float sp[kNP];
  load_params(pb, sp);
  const int lane = threadIdx.x & 63;
  const uint64_t wpb = blockDim.x >> 6;
  const uint64_t wave = blockIdx.x * wpb + (threadIdx.x >> 6);
  const uint64_t nwaves = (uint64_t)gridDim.x * wpb;
  float gW1[kIn], gW2[kH], gW3[kOut], gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
#pragma unroll
  for (int i = 0; i < kIn; ++i) gW1[i] = 0.f;
#pragma unroll
  for (int i = 0; i < kH; ++i) gW2[i] = 0.f;
#pragma unroll
  for (int o = 0; o < kOut; ++o) gW3[o] = 0.f;
  const float inv_n = 1.0f / (float)n;
  for (uint64_t s = wave; s < n; s += nwaves) {
    const float *xs = x + s * kIn;
    float h1, h2, z[kOut];
    fwd(sp, xs, lane, h1, h2, z);
    const float wgt = (wts ? wts[s] : 1.0f) * inv_n;
    float mx = z[0];
#pragma unroll
    for (int o = 1; o < kOut; ++o) mx = fmaxf(mx, z[o]);
    float e[kOut], se = 0.f;
#pragma unroll
    for (int o = 0; o < kOut; ++o) {
      e[o] = expf(z[o] - mx);
      se += e[o];
    }
    const int lab = y[s];
    float zl = z[0];
#pragma unroll
    for (int o = 1; o < kOut; ++o) zl = (o == lab) ? z[o] : zl;
    loss += wgt * (logf(se) + mx - zl);
    float dz[kOut];
#pragma unroll
    for (int o = 0; o < kOut; ++o) dz[o] = wgt * (e[o] / se - (o == lab ? 1.f : 0.f));
    // layer 3: lane j owns column j of W3
    float dh2 = 0.f;
#pragma unroll
    for (int o = 0; o < kOut; ++o) {
      gW3[o] = fmaf(dz[o], h2, gW3[o]);
      dh2 = fmaf(dz[o], sp[oW3 + o * kH + lane], dh2);
    }
    gb3 += lane == 0 ? dz[0] : (lane == 1 ? dz[1] : (lane == 2 ? dz[2] : 0.f));
    dh2 = h2 > 0.f ? dh2 : 0.f;
    // layer 2: lane j owns row j of W2; dh1_i = sum_j dh2_j W2[j][i] is gathered on lane i
    gb2 += dh2;
    float dh1 = 0.f;
#pragma unroll 8
    for (int i = 0; i < kH; ++i) {
      gW2[i] = fmaf(dh2, __shfl(h1, i, 64), gW2[i]);
      dh1 = fmaf(__shfl(dh2, i, 64), sp[oW2 + i * kH + lane], dh1);
    }
    dh1 = h1 > 0.f ? dh1 : 0.f;
    // layer 1: lane j owns row j of W1
    gb1 += dh1;
#pragma unroll
    for (int i = 0; i < kIn; ++i) gW1[i] = fmaf(dh1, xs[i], gW1[i]);
  }
  float *out = partial + wave * (kNP + 1);
#pragma unroll
  for (int i = 0; i < kIn; ++i) out[oW1 + lane * kIn + i] = gW1[i];
  out[oB1 + lane] = gb1;
#pragma unroll 8
  for (int i = 0; i < kH; ++i) out[oW2 + lane * kH + i] = gW2[i];
  out[oB2 + lane] = gb2;
#pragma unroll
  for (int o = 0; o < kOut; ++o) out[oW3 + o * kH + lane] = gW3[o];
  if (lane < kOut) out[oB3 + lane] = gb3;
  // every lane holds the same (wave-uniform) loss
  if (lane == 0) out[kNP] = loss;
}

// grads[k] = sum over waves of partial[w][k], in wave order; k = kNP is the loss.
__global__ void k_mlp_reduce(const float *__restrict__ partial, uint32_t nwaves,
                             float *__restrict__ grads) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > (uint32_t)kNP) return;
  float s = 0.f;
  for (uint32_t w = 0; w < nwaves; ++w) s += partial[(uint64_t)w * (kNP + 1) + k];
  grads[k] = s;
}

// Adam (Kingma & Ba 2015) on the fp32 master weights, then the bf16 copy.
__global__ void k_adam(float *__restrict__ master, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, float lr, float b1t, float b2t,
                       uint16_t *__restrict__ pb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNP) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1.f - b1) * gk;
  const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  const float mh = mk / (1.f - b1t), vh = vk / (1.f - b2t);
  const float w = master[k] - lr * mh / (sqrtf(vh) + eps);
  master[k] = w;
  pb[k] = f32_to_bf16(w);
}

__global__ void k_to_bf16(const float *__restrict__ master, uint16_t *__restrict__ pb) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < kNP) pb[k] = f32_to_bf16(master[k]);
}

constexpr uint32_t kBackwardBlocks = 64;  // x 4 waves: fixed partial count

int ensure_params(Ctx *ctx) {
  PAC_TRY(ctx->mlp_params.reserve(sizeof(uint16_t) * kNP));
  return PAC_OK;
}

int upload_params(Ctx *ctx, const float *params) {
  PAC_TRY(ensure_params(ctx));
  std::vector<uint16_t> b(kNP);
  ctx->mlp_host.resize(kNP);
  for (int i = 0; i < kNP; ++i) {
    b[i] = f32_to_bf16(params[i]);
    ctx->mlp_host[i] = bf16_to_f32(b[i]);
  }
  PAC_CHECK_HIP(hipMemcpy(ctx->mlp_params.p, b.data(), sizeof(uint16_t) * kNP,
                          hipMemcpyHostToDevice));
  ctx->mlp_loaded = true;
  ctx->mlp_version++;
  return PAC_OK;
}

}  // namespace

int mlp_forward_dev(Ctx *ctx, const float *d_x, uint64_t n, float *d_logits) {
  if (n == 0) return PAC_OK;
  uint64_t blocks = (n + 3) / 4;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_mlp_forward, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream,
                     ctx->mlp_params.as<const uint16_t>(), d_x, n, d_logits, nullptr);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int mlp_loss_grad_dev(Ctx *ctx, const float *d_x, const int32_t *d_labels, const float *d_w,
                      uint64_t n, float *d_loss, float *d_grads) {
  // d_grads has kNP + 1 floats: gradient then loss; d_loss unused (kept for clarity)
  (void)d_loss;
  const uint32_t nwaves = kBackwardBlocks * 4;
  PAC_TRY(ctx->scratch2.reserve(sizeof(float) * (size_t)nwaves * (kNP + 1)));
  hipLaunchKernelGGL(k_mlp_backward, dim3(kBackwardBlocks), dim3(256), 0, ctx->stream,
                     ctx->mlp_params.as<const uint16_t>(), d_x, d_labels, d_w, n,
                     ctx->scratch2.as<float>());
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNP + 1 + 255) / 256), dim3(256), 0, ctx->stream,
                     ctx->scratch2.as<const float>(), nwaves, d_grads);
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int mlp_adam_step(Ctx *ctx, float *d_master, float *d_m, float *d_v, const float *d_g, float lr,
                  int step) {
  const float b1t = powf(0.9f, (float)step), b2t = powf(0.999f, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((kNP + 255) / 256), dim3(256), 0, ctx->stream, d_master, d_m,
                     d_v, d_g, lr, b1t, b2t, ctx->mlp_params.as<uint16_t>());
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

int mlp_store_bf16(Ctx *ctx, const float *d_master) {
  hipLaunchKernelGGL(k_to_bf16, dim3((kNP + 255) / 256), dim3(256), 0, ctx->stream, d_master,
                     ctx->mlp_params.as<uint16_t>());
  PAC_CHECK_HIP(hipGetLastError());
  return PAC_OK;
}

}  // namespace pac

using namespace pac;

extern "C" {

int pac_mlp_set_params(pac_ctx *ctx_, const float *params) {
  if (!ctx_ || !params) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  Ctx *ctx = ctx_ ? &ctx_->c : nullptr;
  PAC_CHECK_HIP(hipSetDevice(ctx->device));
  return upload_params(ctx, params);
}

int pac_mlp_get_params(const pac_ctx *ctx_, float *params) {
  const Ctx *ctx = ctx_ ? &ctx_->c : nullptr;
  if (!ctx || !params) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  if (!ctx->mlp_loaded) {
    set_error("no MLP parameters loaded");
    return PAC_ERR_STATE;
  }
  memcpy(params, ctx->mlp_host.data(), sizeof(float) * kNP);
  return PAC_OK;
}

// He-uniform initialisation from a seeded SplitMix64 stream; biases zero.
int pac_mlp_init(pac_ctx *ctx_, uint64_t seed) {
  if (!ctx_) {
    set_error("null context");
    return PAC_ERR_ARG;
  }
  std::vector<float> p(kNP, 0.f);
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
  auto uni = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  const double a1 = std::sqrt(6.0 / kIn), a2 = std::sqrt(6.0 / kH);
  for (int i = 0; i < kH * kIn; ++i) p[oW1 + i] = (float)(a1 * uni());
  for (int i = 0; i < kH * kH; ++i) p[oW2 + i] = (float)(a2 * uni());
  for (int i = 0; i < kOut * kH; ++i) p[oW3 + i] = (float)(a2 * uni() * 0.1);
  return pac_mlp_set_params(ctx_, p.data());
}

int pac_mlp_forward(pac_ctx *ctx_, const float *x, uint64_t n, float *logits) {
  Ctx *ctx = ctx_ ? &ctx_->c : nullptr;
  if (!ctx || (n && (!x || !logits))) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  if (!ctx->mlp_loaded) {
    set_error("no MLP parameters loaded");
    return PAC_ERR_STATE;
  }
  if (n == 0) return PAC_OK;
  PAC_CHECK_HIP(hipSetDevice(ctx->device));
  DevBuf dx, dz;
  PAC_TRY(dx.reserve(sizeof(float) * n * kIn));
  PAC_TRY(dz.reserve(sizeof(float) * n * kOut));
  PAC_CHECK_HIP(hipMemcpy(dx.p, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice));
  PAC_TRY(mlp_forward_dev(ctx, dx.as<float>(), n, dz.as<float>()));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  PAC_CHECK_HIP(hipMemcpy(logits, dz.p, sizeof(float) * n * kOut, hipMemcpyDeviceToHost));
  return PAC_OK;
}

int pac_mlp_loss_grad(pac_ctx *ctx_, const float *x, const int32_t *labels,
                      const float *weights, uint64_t n, float *loss, float *grads) {
  Ctx *ctx = ctx_ ? &ctx_->c : nullptr;
  if (!ctx || !x || !labels || !grads || n == 0) {
    set_error("need context, samples, labels and a gradient buffer");
    return PAC_ERR_ARG;
  }
  if (!ctx->mlp_loaded) {
    set_error("no MLP parameters loaded");
    return PAC_ERR_STATE;
  }
  for (uint64_t i = 0; i < n; ++i)
    if (labels[i] < 0 || labels[i] >= kOut) {
      set_error("label %d out of range", labels[i]);
      return PAC_ERR_ARG;
    }
  PAC_CHECK_HIP(hipSetDevice(ctx->device));
  DevBuf dx, dy, dw, dg;
  PAC_TRY(dx.reserve(sizeof(float) * n * kIn));
  PAC_TRY(dy.reserve(sizeof(int32_t) * n));
  PAC_TRY(dg.reserve(sizeof(float) * (kNP + 1)));
  PAC_CHECK_HIP(hipMemcpy(dx.p, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice));
  PAC_CHECK_HIP(hipMemcpy(dy.p, labels, sizeof(int32_t) * n, hipMemcpyHostToDevice));
  if (weights) {
    PAC_TRY(dw.reserve(sizeof(float) * n));
    PAC_CHECK_HIP(hipMemcpy(dw.p, weights, sizeof(float) * n, hipMemcpyHostToDevice));
  }
  PAC_TRY(mlp_loss_grad_dev(ctx, dx.as<float>(), dy.as<int32_t>(), dw.as<float>(), n, nullptr,
                            dg.as<float>()));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<float> h(kNP + 1);
  PAC_CHECK_HIP(hipMemcpy(h.data(), dg.p, sizeof(float) * (kNP + 1), hipMemcpyDeviceToHost));
  memcpy(grads, h.data(), sizeof(float) * kNP);
  if (loss) *loss = h[kNP];
  return PAC_OK;
}

int pac_mlp_train(pac_ctx *ctx_, const float *x, const int32_t *labels, const float *weights,
                  uint64_t n, uint32_t epochs, float lr, float *loss_out) {
  Ctx *ctx = ctx_ ? &ctx_->c : nullptr;
  if (!ctx || !x || !labels || n == 0) {
    set_error("need context, samples and labels");
    return PAC_ERR_ARG;
  }
  if (!ctx->mlp_loaded) PAC_TRY(pac_mlp_init(ctx_, 1));
  for (uint64_t i = 0; i < n; ++i)
    if (labels[i] < 0 || labels[i] >= kOut) {
      set_error("label %d out of range", labels[i]);
      return PAC_ERR_ARG;
    }
  PAC_CHECK_HIP(hipSetDevice(ctx->device));
  DevBuf dx, dy, dw, dg, dmaster, dm, dv;
  PAC_TRY(dx.reserve(sizeof(float) * n * kIn));
  PAC_TRY(dy.reserve(sizeof(int32_t) * n));
  PAC_TRY(dg.reserve(sizeof(float) * (kNP + 1)));
  PAC_TRY(dmaster.reserve(sizeof(float) * kNP));
  PAC_TRY(dm.reserve(sizeof(float) * kNP));
  PAC_TRY(dv.reserve(sizeof(float) * kNP));
  PAC_CHECK_HIP(hipMemcpy(dx.p, x, sizeof(float) * n * kIn, hipMemcpyHostToDevice));
  PAC_CHECK_HIP(hipMemcpy(dy.p, labels, sizeof(int32_t) * n, hipMemcpyHostToDevice));
  if (weights) {
    PAC_TRY(dw.reserve(sizeof(float) * n));
    PAC_CHECK_HIP(hipMemcpy(dw.p, weights, sizeof(float) * n, hipMemcpyHostToDevice));
  }
  PAC_CHECK_HIP(hipMemcpy(dmaster.p, ctx->mlp_host.data(), sizeof(float) * kNP,
                          hipMemcpyHostToDevice));
  PAC_TRY(fill_u32(ctx->stream, dm.as<uint32_t>(), kNP, 0u));
  PAC_TRY(fill_u32(ctx->stream, dv.as<uint32_t>(), kNP, 0u));
  PAC_TRY(mlp_store_bf16(ctx, dmaster.as<float>()));
  for (uint32_t e = 1; e <= epochs; ++e) {
    PAC_TRY(mlp_loss_grad_dev(ctx, dx.as<float>(), dy.as<int32_t>(), dw.as<float>(), n, nullptr,
                              dg.as<float>()));
    PAC_TRY(mlp_adam_step(ctx, dmaster.as<float>(), dm.as<float>(), dv.as<float>(),
                          dg.as<float>(), lr, (int)e));
  }
  // final loss with the trained weights
  PAC_TRY(mlp_loss_grad_dev(ctx, dx.as<float>(), dy.as<int32_t>(), dw.as<float>(), n, nullptr,
                            dg.as<float>()));
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::vector<uint16_t> b(kNP);
  PAC_CHECK_HIP(hipMemcpy(b.data(), ctx->mlp_params.p, sizeof(uint16_t) * kNP,
                          hipMemcpyDeviceToHost));
  for (int i = 0; i < kNP; ++i) ctx->mlp_host[i] = bf16_to_f32(b[i]);
  ctx->mlp_loaded = true;
  ctx->mlp_version++;
  if (loss_out)
    PAC_CHECK_HIP(hipMemcpy(loss_out, dg.as<float>() + kNP, sizeof(float), hipMemcpyDeviceToHost));
  return PAC_OK;
}

int pac_mlp_save(const pac_ctx *ctx_, const char *path) {
  const Ctx *ctx = ctx_ ? &ctx_->c : nullptr;
  if (!ctx || !path) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  if (!ctx->mlp_loaded) {
    set_error("no MLP parameters loaded");
    return PAC_ERR_STATE;
  }
  FILE *f = fopen(path, "wb");
  if (!f) {
    set_error("cannot create %s", path);
    return PAC_ERR_IO;
  }
  const char magic[8] = {'P', 'A', 'C', 'M', 'L', 'P', '1', 0};
  const uint32_t np = kNP;
  bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(&np, 4, 1, f) == 1 &&
            fwrite(ctx->mlp_host.data(), sizeof(float), kNP, f) == (size_t)kNP;
  ok = (fclose(f) == 0) && ok;
  if (!ok) {
    set_error("%s: write failed", path);
    return PAC_ERR_IO;
  }
  return PAC_OK;
}

int pac_mlp_load(pac_ctx *ctx_, const char *path) {
  if (!ctx_ || !path) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  FILE *f = fopen(path, "rb");
  if (!f) {
    set_error("cannot open %s", path);
    return PAC_ERR_IO;
  }
  char magic[8];
  uint32_t np = 0;
  std::vector<float> p(kNP);
  const bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "PACMLP1", 7) == 0 &&
                  fread(&np, 4, 1, f) == 1 && np == (uint32_t)kNP &&
                  fread(p.data(), sizeof(float), kNP, f) == (size_t)kNP;
  fclose(f);
  if (!ok) {
    set_error("%s: not a PACMLP1 parameter file", path);
    return PAC_ERR_FORMAT;
  }
  return pac_mlp_set_params(ctx_, p.data());
}

int pac_mlp_choice(pac_graph *g_, uint8_t *choice) {
  Graph *g = g_ ? &g_->g : nullptr;
  if (!g || !choice) {
    set_error("null argument");
    return PAC_ERR_ARG;
  }
  if (!g->prepared) {
    set_error("graph not prepared");
    return PAC_ERR_STATE;
  }
  Ctx *ctx = g->ctx;
  if (!ctx->mlp_loaded) {
    set_error("no MLP parameters loaded");
    return PAC_ERR_STATE;
  }
  PAC_CHECK_HIP(hipSetDevice(ctx->device));
  if (g->P == 0) return PAC_OK;
  DevBuf dx, dc;
  PAC_TRY(dx.reserve(sizeof(float) * g->h_features.size()));
  PAC_TRY(dc.reserve(g->P));
  PAC_CHECK_HIP(hipMemcpy(dx.p, g->h_features.data(), sizeof(float) * g->h_features.size(),
                          hipMemcpyHostToDevice));
  uint64_t blocks = (g->P + 3) / 4;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_mlp_forward, dim3((uint32_t)blocks), dim3(256), 0, ctx->stream,
                     ctx->mlp_params.as<const uint16_t>(), dx.as<const float>(), (uint64_t)g->P,
                     nullptr, dc.as<uint8_t>());
  PAC_CHECK_HIP(hipGetLastError());
  PAC_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  PAC_CHECK_HIP(hipMemcpy(choice, dc.p, g->P, hipMemcpyDeviceToHost));
  return PAC_OK;
}

}  // extern "C"
