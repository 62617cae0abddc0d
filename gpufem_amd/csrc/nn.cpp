// Learned warm start: the network (a four-level 3D U-Net on the node grid),
// its training step (forward, MSE, backward, Adam) and the C-ABI around it.
//
// Levels l = 0..3 are vertex-centred node grids, n_{l+1} = (n_l + 1) / 2 per
// axis (so every axis length is 1 mod 8).  Down-sampling is a stride-2
// convolution centred on the even fine nodes, up-sampling a trilinear
// prolongation; skip connections are concatenations realised by writing the
// encoder output straight into the upper channels of the decoder's input.
//
//   e0a 16->16, e0b 16->16 (skip0)          level 0
//   e1a 16->32 s2, e1b 32->32 (skip1)       level 1
//   e2a 32->64 s2, e2b 64->64 (skip2)       level 2
//   e3a 64->128 s2                          level 3
//   up, [128 | skip2] d2a 192->64, d2b      level 2
//   up, [64 | skip1]  d1a 96->32, d1b       level 1
//   up, [32 | skip0]  d0a 48->16, d0b       level 0
//   out 16->16 (FP32, no activation; the first 3 channels are displacement)
// Every hidden convolution has a bias and a leaky ReLU (slope 0.1).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gf_internal.hpp"
#include "nn.hpp"

using gf::nn::Vol;

namespace {

constexpr int kCin = 16;     // input channels (features, zero-padded)
constexpr int kCout = 16;    // output channels (first kOutUsed are used)
constexpr int kOutUsed = 3;
constexpr float kSlope = 0.1f;

enum Buf {
  B_IN = 0, B_A0, B_CAT0, B_A1, B_CAT1, B_A2, B_CAT2, B_A3, B_B2, B_C2, B_B1, B_C1, B_B0, B_C0,
  B_PRED, B_COUNT
};

struct Tensor {
  int buf, ld, off, level;
};

struct Layer {
  const char* name;
  Tensor in, out;
  int cin, cout, stride, act, accumulate_dx, need_dx;
  int64_t w_off = 0, b_off = 0;  // offsets in the flat parameter vector
  uint16_t* wf = nullptr;        // forward fragments
  uint16_t* wt = nullptr;        // transposed fragments (data gradient)
};

struct Up {
  Tensor coarse, fine;
  int c;
};

}  // namespace

struct gf_nn {
  int dev = 0;
  Vol lv[4];
  int batch_max = 0;
  std::vector<Layer> layers;
  Up ups[3];
  int64_t nparams = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // fp32 [nparams]
  int64_t adam_t = 0;
  uint16_t* act[B_COUNT] = {};   // activations (B_IN unused: caller's tensor)
  uint16_t* grad[B_COUNT] = {};  // their gradients (B_PRED: dpred)
  float* pred = nullptr;         // fp32 [batch][V0][16]
  float* wpart = nullptr;        // weight-gradient partials
  int64_t wpart_len = 0;
  float* spart = nullptr;        // small partials (bias, loss)
  float* loss = nullptr;
  std::vector<void*> allocs;
};

namespace {

using gf::set_error;
using namespace gf::nn;

int ld_of(int buf) {
  switch (buf) {
    case B_IN: return kCin;
    case B_A0: case B_B0: case B_C0: return 16;
    case B_CAT0: return 48;
    case B_A1: case B_B1: case B_C1: return 32;
    case B_CAT1: return 96;
    case B_A2: case B_B2: case B_C2: return 64;
    case B_CAT2: return 192;
    case B_A3: return 128;
    case B_PRED: return kCout;
    default: return 0;
  }
}

int level_of(int buf) {
  switch (buf) {
    case B_IN: case B_A0: case B_CAT0: case B_B0: case B_C0: case B_PRED: return 0;
    case B_A1: case B_CAT1: case B_B1: case B_C1: return 1;
    case B_A2: case B_CAT2: case B_B2: case B_C2: return 2;
    default: return 3;
  }
}

Tensor T(int buf, int off) { return Tensor{buf, ld_of(buf), off, level_of(buf)}; }

void build_layers(gf_nn* net) {
  auto L = [&](const char* name, Tensor in, Tensor out, int cin, int cout, int stride, int act,
               int acc, int need_dx) {
    Layer l;
    l.name = name;
    l.in = in;
    l.out = out;
    l.cin = cin;
    l.cout = cout;
    l.stride = stride;
    l.act = act;
    l.accumulate_dx = acc;
    l.need_dx = need_dx;
    net->layers.push_back(l);
  };
  L("e0a", T(B_IN, 0), T(B_A0, 0), 16, 16, 1, 1, 0, 0);
  L("e0b", T(B_A0, 0), T(B_CAT0, 32), 16, 16, 1, 1, 0, 1);
  L("e1a", T(B_CAT0, 32), T(B_A1, 0), 16, 32, 2, 1, 1, 1);
  L("e1b", T(B_A1, 0), T(B_CAT1, 64), 32, 32, 1, 1, 0, 1);
  L("e2a", T(B_CAT1, 64), T(B_A2, 0), 32, 64, 2, 1, 1, 1);
  L("e2b", T(B_A2, 0), T(B_CAT2, 128), 64, 64, 1, 1, 0, 1);
  L("e3a", T(B_CAT2, 128), T(B_A3, 0), 64, 128, 2, 1, 1, 1);
  L("d2a", T(B_CAT2, 0), T(B_B2, 0), 192, 64, 1, 1, 0, 1);
  L("d2b", T(B_B2, 0), T(B_C2, 0), 64, 64, 1, 1, 0, 1);
  L("d1a", T(B_CAT1, 0), T(B_B1, 0), 96, 32, 1, 1, 0, 1);
  L("d1b", T(B_B1, 0), T(B_C1, 0), 32, 32, 1, 1, 0, 1);
  L("d0a", T(B_CAT0, 0), T(B_B0, 0), 48, 16, 1, 1, 0, 1);
  L("d0b", T(B_B0, 0), T(B_C0, 0), 16, 16, 1, 1, 0, 1);
  L("out", T(B_C0, 0), T(B_PRED, 0), 16, kCout, 1, 0, 0, 1);
  net->ups[0] = Up{T(B_A3, 0), T(B_CAT2, 0), 128};  // before d2a
  net->ups[1] = Up{T(B_C2, 0), T(B_CAT1, 0), 64};   // before d1a
  net->ups[2] = Up{T(B_C1, 0), T(B_CAT0, 0), 32};   // before d0a
  int64_t off = 0;
  for (Layer& l : net->layers) {
    l.w_off = off;
    off += int64_t(27) * l.cin * l.cout;
    l.b_off = off;
    off += l.cout;
  }
  net->nparams = off;
}

// layer index after which prolongation u runs (forward)
int up_after(int u) { return u == 0 ? 6 : (u == 1 ? 8 : 10); }

template <class P>
hipError_t dalloc(gf_nn* net, P** p, size_t bytes) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e == hipSuccess) {
    net->allocs.push_back(q);
    *p = static_cast<P*>(q);
  }
  return e;
}

void release(gf_nn* net) {
  for (void* q : net->allocs) (void)hipFree(q);
  net->allocs.clear();
}

int64_t voxels(const gf_nn* net, int level, int batch) { return int64_t(batch) * net->lv[level].nodes(); }

hipError_t arrange_all(hipStream_t s, gf_nn* net) {
  for (Layer& l : net->layers) {
    hipError_t e = launch_arrange(s, net->p + l.w_off, l.wf, l.cin, l.cout, 0);
    if (e == hipSuccess && l.need_dx) e = launch_arrange(s, net->p + l.w_off, l.wt, l.cin, l.cout, 1);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

const uint16_t* tensor_ptr(const gf_nn* net, const Tensor& t, const uint16_t* input) {
  return t.buf == B_IN ? input : net->act[t.buf];
}

hipError_t run_layer(hipStream_t s, gf_nn* net, const Layer& l, const uint16_t* input, int batch) {
  ConvArgs a{};
  a.in = tensor_ptr(net, l.in, input);
  a.in_ld = l.in.ld;
  a.in_off = l.in.off;
  if (l.out.buf == B_PRED) {
    a.out32 = net->pred;
  } else {
    a.out = net->act[l.out.buf];
  }
  a.out_ld = l.out.ld;
  a.out_off = l.out.off;
  a.wf = l.wf;
  a.bias = net->p + l.b_off;
  a.batch = batch;
  a.vin = net->lv[l.in.level];
  a.vout = net->lv[l.out.level];
  a.cin = l.cin;
  a.cout = l.cout;
  a.stride = l.stride;
  a.transposed = 0;
  a.act = l.act;
  a.accumulate = 0;
  a.slope = kSlope;
  return launch_conv(s, a);
}

hipError_t forward(hipStream_t s, gf_nn* net, const uint16_t* input, int batch) {
  for (size_t i = 0; i < net->layers.size(); ++i) {
    const Layer& l = net->layers[i];
    hipError_t e = run_layer(s, net, l, input, batch);
    if (e != hipSuccess) return e;
    for (int u = 0; u < 3; ++u) {
      if (up_after(u) != int(i)) continue;
      const Up& up = net->ups[u];
      e = launch_prolong(s, net->act[up.coarse.buf], up.coarse.ld, up.coarse.off,
                         net->lv[up.coarse.level], net->act[up.fine.buf], up.fine.ld, up.fine.off,
                         net->lv[up.fine.level], batch, up.c);
      if (e != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

hipError_t backward(hipStream_t s, gf_nn* net, const uint16_t* input, int batch) {
  for (int i = int(net->layers.size()) - 1; i >= 0; --i) {
    // prolongations that ran after layer i: their adjoints come first
    for (int u = 2; u >= 0; --u) {
      if (up_after(u) != i) continue;
      const Up& up = net->ups[u];
      hipError_t e = launch_prolong_bwd(s, net->grad[up.fine.buf], up.fine.ld, up.fine.off,
                                        net->lv[up.fine.level], net->grad[up.coarse.buf],
                                        up.coarse.ld, up.coarse.off, net->lv[up.coarse.level],
                                        batch, up.c);
      if (e != hipSuccess) return e;
    }
    const Layer& l = net->layers[size_t(i)];
    uint16_t* dz = net->grad[l.out.buf];
    const int64_t nvo = voxels(net, l.out.level, batch);
    hipError_t e = hipSuccess;
    if (l.act)
      e = launch_act_bwd(s, dz, l.out.ld, l.out.off, net->act[l.out.buf], l.out.ld, l.out.off, nvo,
                         l.cout, kSlope);
    if (e != hipSuccess) return e;
    WgradArgs w{};
    w.x = tensor_ptr(net, l.in, input);
    w.x_ld = l.in.ld;
    w.x_off = l.in.off;
    w.dz = dz;
    w.dz_ld = l.out.ld;
    w.dz_off = l.out.off;
    w.part = net->wpart;
    w.batch = batch;
    w.vin = net->lv[l.in.level];
    w.vout = net->lv[l.out.level];
    w.cin = l.cin;
    w.cout = l.cout;
    w.stride = l.stride;
    w.nchunk = wgrad_chunks(nvo);
    w.chunk = ((nvo + w.nchunk - 1) / w.nchunk + 63) / 64 * 64;
    if (int64_t(w.nchunk) * 27 * l.cin * l.cout > net->wpart_len) return hipErrorInvalidValue;
    e = launch_wgrad(s, w, net->g + l.w_off);
    if (e != hipSuccess) return e;
    e = launch_bias_grad(s, dz, l.out.ld, l.out.off, nvo, l.cout, net->spart, net->g + l.b_off);
    if (e != hipSuccess) return e;
    if (!l.need_dx) continue;
    ConvArgs a{};
    a.in = dz;
    a.in_ld = l.out.ld;
    a.in_off = l.out.off;
    a.out = net->grad[l.in.buf];
    a.out_ld = l.in.ld;
    a.out_off = l.in.off;
    a.wf = l.wt;
    a.bias = nullptr;
    a.batch = batch;
    a.vin = net->lv[l.out.level];
    a.vout = net->lv[l.in.level];
    a.cin = l.cout;
    a.cout = l.cin;
    a.stride = l.stride;
    a.transposed = 1;
    a.act = 0;
    a.accumulate = l.accumulate_dx;
    a.slope = 0.f;
    e = launch_conv(s, a);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

int check_batch(gf_ctx* ctx, gf_nn* net, int batch, const char* fn) {
  if (batch < 1 || batch > net->batch_max)
    return set_error(ctx, GF_ERR_ARG, std::string(fn) + ": batch outside 1..batch_max");
  return GF_OK;
}

}  // namespace

extern "C" {

int gf_nn_create(gf_ctx* ctx, int32_t nx, int32_t ny, int32_t nz, int32_t batch_max,
                 uint64_t seed, gf_nn** out) {
  if (!ctx || !out || batch_max < 1) return set_error(ctx, GF_ERR_ARG, "gf_nn_create: bad argument");
  for (int n : {nx, ny, nz})
    if (n < 9 || (n - 1) % 8 != 0)
      return set_error(ctx, GF_ERR_ARG, "gf_nn_create: every node-grid axis must be 1 mod 8 and >= 9");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  gf_nn* net = new gf_nn();
  net->dev = ctx->device;
  net->batch_max = batch_max;
  net->lv[0] = Vol{nx, ny, nz};
  for (int l = 1; l < 4; ++l)
    net->lv[l] = Vol{(net->lv[l - 1].nx + 1) / 2, (net->lv[l - 1].ny + 1) / 2, (net->lv[l - 1].nz + 1) / 2};
  build_layers(net);
  hipError_t e = hipSuccess;
  const size_t pb = size_t(net->nparams) * sizeof(float);
  for (float** q : {&net->p, &net->g, &net->m, &net->v})
    if (e == hipSuccess) e = dalloc(net, q, pb);
  for (int b = B_A0; b < B_COUNT; ++b) {
    const size_t bytes = size_t(voxels(net, level_of(b), batch_max)) * ld_of(b) * sizeof(uint16_t);
    if (e == hipSuccess && b != B_PRED) e = dalloc(net, &net->act[b], bytes);
    if (e == hipSuccess) e = dalloc(net, &net->grad[b], bytes);
  }
  if (e == hipSuccess)
    e = dalloc(net, &net->pred, size_t(voxels(net, 0, batch_max)) * kCout * sizeof(float));
  int64_t wl = 0;
  for (const Layer& l : net->layers)
    wl = std::max<int64_t>(wl, int64_t(wgrad_chunks(voxels(net, l.out.level, batch_max))) * 27 *
                                   l.cin * l.cout);
  net->wpart_len = wl;
  if (e == hipSuccess) e = dalloc(net, &net->wpart, size_t(wl) * sizeof(float));
  if (e == hipSuccess) e = dalloc(net, &net->spart, size_t(256 * 256) * sizeof(float));
  if (e == hipSuccess) e = dalloc(net, &net->loss, sizeof(float));
  for (Layer& l : net->layers) {
    const size_t fb = size_t(27) * l.cin * l.cout * sizeof(uint16_t);
    if (e == hipSuccess) e = dalloc(net, &l.wf, fb);
    if (e == hipSuccess && l.need_dx) e = dalloc(net, &l.wt, fb);
  }
  if (e != hipSuccess) {
    release(net);
    delete net;
    return set_error(ctx, GF_ERR_NOMEM, std::string("gf_nn_create: ") + hipGetErrorString(e));
  }
  // He-normal weights (fan-in 27 cin; the input layer counts its 8 used
  // channels), zero biases; the output layer starts 10x smaller
  std::vector<double> z;
  std::vector<float> h(size_t(net->nparams), 0.f);
  uint64_t sd = seed;
  for (const Layer& l : net->layers) {
    const int64_t nw = int64_t(27) * l.cin * l.cout;
    z.resize(size_t(nw));
    gf_rng_normal(sd++, nw, z.data());
    const int fan = 27 * (l.in.buf == B_IN ? 8 : l.cin);
    const double sdev = std::sqrt(2.0 / fan) * (l.out.buf == B_PRED ? 0.1 : 1.0);
    for (int64_t i = 0; i < nw; ++i) h[size_t(l.w_off + i)] = float(sdev * z[size_t(i)]);
  }
  e = hipMemcpy(net->p, h.data(), pb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(net->m, 0, pb);
  if (e == hipSuccess) e = hipMemset(net->v, 0, pb);
  if (e == hipSuccess) e = hipMemset(net->g, 0, pb);
  if (e == hipSuccess) e = arrange_all(ctx->stream, net);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    release(net);
    delete net;
    return set_error(ctx, GF_ERR_HIP, std::string("gf_nn_create: ") + hipGetErrorString(e));
  }
  *out = net;
  return GF_OK;
}

int gf_nn_free(gf_nn* net) {
  if (!net) return GF_OK;
  (void)hipSetDevice(net->dev);
  release(net);
  delete net;
  return GF_OK;
}

int64_t gf_nn_nparams(const gf_nn* net) { return net ? net->nparams : -1; }

int gf_nn_layer_count(const gf_nn* net) { return net ? int(net->layers.size()) : -1; }

int gf_nn_layer_info(const gf_nn* net, int32_t i, int32_t* info) {
  if (!net || !info || i < 0 || i >= int(net->layers.size())) return GF_ERR_ARG;
  const Layer& l = net->layers[size_t(i)];
  info[0] = l.cin;
  info[1] = l.cout;
  info[2] = l.stride;
  info[3] = l.act;
  info[4] = l.in.level;
  info[5] = l.out.level;
  info[6] = int32_t(l.w_off);
  info[7] = int32_t(l.b_off);
  return GF_OK;
}

int gf_nn_params_get(gf_ctx* ctx, gf_nn* net, float* host) {
  if (!ctx || !net || !host) return set_error(ctx, GF_ERR_ARG, "gf_nn_params_get: null argument");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  GF_HIP_CHECK(ctx, hipMemcpy(host, net->p, size_t(net->nparams) * sizeof(float), hipMemcpyDeviceToHost));
  return GF_OK;
}

int gf_nn_params_set(gf_ctx* ctx, gf_nn* net, const float* host) {
  if (!ctx || !net || !host) return set_error(ctx, GF_ERR_ARG, "gf_nn_params_set: null argument");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  GF_HIP_CHECK(ctx, hipMemcpy(net->p, host, size_t(net->nparams) * sizeof(float), hipMemcpyHostToDevice));
  GF_HIP_CHECK(ctx, arrange_all(ctx->stream, net));
  net->adam_t = 0;
  GF_HIP_CHECK(ctx, hipMemsetAsync(net->m, 0, size_t(net->nparams) * sizeof(float), ctx->stream));
  GF_HIP_CHECK(ctx, hipMemsetAsync(net->v, 0, size_t(net->nparams) * sizeof(float), ctx->stream));
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return GF_OK;
}

int gf_nn_grads_get(gf_ctx* ctx, gf_nn* net, float* host) {
  if (!ctx || !net || !host) return set_error(ctx, GF_ERR_ARG, "gf_nn_grads_get: null argument");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  GF_HIP_CHECK(ctx, hipMemcpy(host, net->g, size_t(net->nparams) * sizeof(float), hipMemcpyDeviceToHost));
  return GF_OK;
}

int gf_nn_forward(gf_ctx* ctx, gf_nn* net, const uint16_t* in_dev, int32_t batch, float* out_dev) {
  if (!ctx || !net || !in_dev) return set_error(ctx, GF_ERR_ARG, "gf_nn_forward: null argument");
  int st = check_batch(ctx, net, batch, "gf_nn_forward");
  if (st) return st;
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  GF_HIP_CHECK(ctx, forward(ctx->stream, net, in_dev, batch));
  if (out_dev)
    GF_HIP_CHECK(ctx, hipMemcpyAsync(out_dev, net->pred,
                                     size_t(voxels(net, 0, batch)) * kCout * sizeof(float),
                                     hipMemcpyDeviceToDevice, ctx->stream));
  return GF_OK;
}

int gf_nn_backward(gf_ctx* ctx, gf_nn* net, const uint16_t* in_dev, const float* target_dev,
                   int32_t batch, float* loss_host) {
  if (!ctx || !net || !in_dev || !target_dev)
    return set_error(ctx, GF_ERR_ARG, "gf_nn_backward: null argument");
  int st = check_batch(ctx, net, batch, "gf_nn_backward");
  if (st) return st;
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  GF_HIP_CHECK(ctx, forward(s, net, in_dev, batch));
  GF_HIP_CHECK(ctx, launch_mse(s, net->pred, target_dev, voxels(net, 0, batch), kOutUsed,
                               net->grad[B_PRED], net->spart, net->loss));
  GF_HIP_CHECK(ctx, backward(s, net, in_dev, batch));
  if (loss_host) {
    GF_HIP_CHECK(ctx, hipMemcpyAsync(loss_host, net->loss, sizeof(float), hipMemcpyDeviceToHost, s));
    GF_HIP_CHECK(ctx, hipStreamSynchronize(s));
  }
  return GF_OK;
}

int gf_nn_adam(gf_ctx* ctx, gf_nn* net, float lr, float beta1, float beta2, float eps) {
  if (!ctx || !net || !(lr >= 0.f)) return set_error(ctx, GF_ERR_ARG, "gf_nn_adam: bad argument");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  ++net->adam_t;
  const double t = double(net->adam_t);
  const float lr_t = float(lr * std::sqrt(1.0 - std::pow(double(beta2), t)) /
                           (1.0 - std::pow(double(beta1), t)));
  GF_HIP_CHECK(ctx, launch_adam(ctx->stream, net->p, net->g, net->m, net->v, net->nparams, lr_t,
                                beta1, beta2, eps));
  GF_HIP_CHECK(ctx, arrange_all(ctx->stream, net));
  return GF_OK;
}

int gf_nn_train_step(gf_ctx* ctx, gf_nn* net, const uint16_t* in_dev, const float* target_dev,
                     int32_t batch, float lr, float* loss_host) {
  int st = gf_nn_backward(ctx, net, in_dev, target_dev, batch, nullptr);
  if (st) return st;
  st = gf_nn_adam(ctx, net, lr, 0.9f, 0.999f, 1e-8f);
  if (st) return st;
  if (loss_host) {
    GF_HIP_CHECK(ctx, hipMemcpyAsync(loss_host, net->loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    GF_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GF_OK;
}

int gf_nn_to_dofs(gf_ctx* ctx, gf_nn* net, const float* pred_dev, const uint16_t* in_dev,
                  int32_t sample, int32_t mask_channel, int32_t base_channel, double base_weight,
                  double scale, double* x_dev) {
  if (!ctx || !net || !pred_dev || !in_dev || !x_dev || sample < 0 || sample >= net->batch_max)
    return set_error(ctx, GF_ERR_ARG, "gf_nn_to_dofs: bad argument (sample outside 0..batch_max-1)");
  if (mask_channel < 0 || mask_channel >= kCin || base_channel < -1 || base_channel > kCin - 3)
    return set_error(ctx, GF_ERR_ARG, "gf_nn_to_dofs: channel outside the input channels");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const int64_t nodes = net->lv[0].nodes();
  GF_HIP_CHECK(ctx, launch_to_dofs(ctx->stream, pred_dev + sample * nodes * kCout,
                                   in_dev + sample * nodes * kCin, mask_channel, base_channel,
                                   base_weight, nodes, scale, x_dev));
  return GF_OK;
}

int gf_nn_conv3d(gf_ctx* ctx, const uint16_t* in_dev, int32_t batch, const int32_t* vin,
                 int32_t cin, const float* w_dev, const float* bias_dev, int32_t cout,
                 int32_t stride, int32_t transposed, int32_t act, float slope, const int32_t* vout,
                 uint16_t* out_dev, float* out32_dev, int32_t accumulate) {
  if (!ctx || !in_dev || !vin || !vout || !w_dev || (!out_dev == !out32_dev))
    return set_error(ctx, GF_ERR_ARG, "gf_nn_conv3d: bad argument");
  if ((cin & 15) || (cout & 15) || cin < 16 || cout < 16 || (stride != 1 && stride != 2))
    return set_error(ctx, GF_ERR_ARG, "gf_nn_conv3d: channels must be multiples of 16, stride 1 or 2");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // fragments of the weights W[27][cin][cout] (or, transposed, W[27][cout][cin])
  uint16_t* wf = nullptr;
  GF_HIP_CHECK(ctx, hipMalloc(&wf, size_t(27) * cin * cout * sizeof(uint16_t)));
  hipError_t e = transposed ? launch_arrange(ctx->stream, w_dev, wf, cout, cin, 1)
                            : launch_arrange(ctx->stream, w_dev, wf, cin, cout, 0);
  ConvArgs a{};
  a.in = in_dev;
  a.in_ld = cin;
  a.in_off = 0;
  a.out = out_dev;
  a.out32 = out32_dev;
  a.out_ld = cout;
  a.out_off = 0;
  a.wf = wf;
  a.bias = bias_dev;
  a.batch = batch;
  a.vin = Vol{vin[0], vin[1], vin[2]};
  a.vout = Vol{vout[0], vout[1], vout[2]};
  a.cin = cin;
  a.cout = cout;
  a.stride = stride;
  a.transposed = transposed;
  a.act = act;
  a.accumulate = accumulate;
  a.slope = slope;
  if (e == hipSuccess) e = launch_conv(ctx->stream, a);
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  (void)hipFree(wf);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return set_error(ctx, GF_ERR_HIP, std::string("gf_nn_conv3d: ") + hipGetErrorString(e));
  return GF_OK;
}

int gf_nn_conv3d_wgrad(gf_ctx* ctx, const uint16_t* x_dev, int32_t batch, const int32_t* vin,
                       int32_t cin, const uint16_t* dz_dev, const int32_t* vout, int32_t cout,
                       int32_t stride, float* dw_dev, float* db_dev) {
  if (!ctx || !x_dev || !dz_dev || !vin || !vout || !dw_dev)
    return set_error(ctx, GF_ERR_ARG, "gf_nn_conv3d_wgrad: bad argument");
  if ((cin & 15) || (cout & 15) || cin < 16 || cout < 16 || cout > 256 || (stride != 1 && stride != 2))
    return set_error(ctx, GF_ERR_ARG, "gf_nn_conv3d_wgrad: bad channels or stride");
  GF_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  WgradArgs w{};
  w.x = x_dev;
  w.x_ld = cin;
  w.x_off = 0;
  w.dz = dz_dev;
  w.dz_ld = cout;
  w.dz_off = 0;
  w.batch = batch;
  w.vin = Vol{vin[0], vin[1], vin[2]};
  w.vout = Vol{vout[0], vout[1], vout[2]};
  w.cin = cin;
  w.cout = cout;
  w.stride = stride;
  const int64_t nvo = int64_t(batch) * w.vout.nodes();
  w.nchunk = wgrad_chunks(nvo);
  w.chunk = ((nvo + w.nchunk - 1) / w.nchunk + 63) / 64 * 64;
  float* part = nullptr;
  GF_HIP_CHECK(ctx, hipMalloc(&part, (size_t(w.nchunk) * 27 * cin * cout + 256 * 256) * sizeof(float)));
  w.part = part;
  hipError_t e = launch_wgrad(ctx->stream, w, dw_dev);
  if (e == hipSuccess && db_dev)
    e = launch_bias_grad(ctx->stream, dz_dev, cout, 0, nvo, cout,
                         part + size_t(w.nchunk) * 27 * cin * cout, db_dev);
  hipError_t e2 = hipStreamSynchronize(ctx->stream);
  (void)hipFree(part);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess)
    return set_error(ctx, GF_ERR_HIP, std::string("gf_nn_conv3d_wgrad: ") + hipGetErrorString(e));
  return GF_OK;
}

}  // extern "C"
