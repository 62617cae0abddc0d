// Learned warm start: the layer kernels of a small BF16 3D convolutional
// encoder-decoder on a vertex-centred node grid (kernels in nn.hip, network
// and training loop in nn.cpp).
//
// Activations are channels-last: element (b, z, y, x, c) of a tensor with
// row stride `ld` (channels, a multiple of 16) lives at
// ((b*nz + z)*ny + y)*nx + x)*ld + c, stored as bfloat16 bit patterns.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gf {
namespace nn {

struct Vol {
  int nx, ny, nz;
  __host__ __device__ int64_t nodes() const { return int64_t(nx) * ny * nz; }
};

// 3x3x3 convolution (zero padding) as an implicit GEMM on MFMA.
//   forward    (transposed = 0): out[v] = sum_t in[s*v + t] F_t   (t in {-1,0,1}^3)
//   transposed (transposed = 1): out[u] = sum_t in[(u - t)/s] F_t  (divisible, in range)
// F_t[c][o] is taken from `wf`, the weights arranged as MFMA A fragments by
// launch_arrange; bias, leaky ReLU and accumulation into `out` are fused into
// the epilogue.  Exactly one of out / out32 is set.
struct ConvArgs {
  const uint16_t* in;
  int in_ld, in_off;
  uint16_t* out;
  float* out32;
  int out_ld, out_off;
  const uint16_t* wf;
  const float* bias;
  int batch;
  Vol vin, vout;
  int cin, cout;  // channels the kernel reads / writes (multiples of 16)
  int stride, transposed, act, accumulate;
  float slope;
};

// Weight gradient of a forward convolution:
//   dW[t][ci][co] = sum_v x[s*v + t][ci] * dz[v][co], per chunk of voxels
struct WgradArgs {
  const uint16_t* x;
  int x_ld, x_off;
  const uint16_t* dz;
  int dz_ld, dz_off;
  float* part;  // [nchunk][27][cin][cout]
  int batch;
  Vol vin, vout;
  int cin, cout, stride;
  int64_t chunk;  // voxels per chunk (multiple of 64)
  int nchunk;
};

hipError_t launch_conv(hipStream_t s, const ConvArgs& a);
int wgrad_chunks(int64_t voxels);
hipError_t launch_wgrad(hipStream_t s, const WgradArgs& a, float* dw);
// fp32 master weights W[27][cin][cout] -> MFMA fragments; transposed
// swaps the roles of input and output channels (for the data gradient)
hipError_t launch_arrange(hipStream_t s, const float* w, uint16_t* wf, int cin, int cout,
                          int transposed);
// y > 0 ? dy : slope * dy, in place on the channel slice [off, off + c)
hipError_t launch_act_bwd(hipStream_t s, uint16_t* dy, int dy_ld, int dy_off, const uint16_t* y,
                          int y_ld, int y_off, int64_t voxels, int c, float slope);
// db[c] = sum over voxels of dz[v][c]
hipError_t launch_bias_grad(hipStream_t s, const uint16_t* dz, int ld, int off, int64_t voxels,
                            int c, float* part, float* db);
// vertex-centred trilinear prolongation coarse (vc) -> fine (vf = 2 vc - 1)
// and its adjoint; both overwrite their output slice
hipError_t launch_prolong(hipStream_t s, const uint16_t* in, int in_ld, int in_off, Vol vc,
                          uint16_t* out, int out_ld, int out_off, Vol vf, int batch, int c);
hipError_t launch_prolong_bwd(hipStream_t s, const uint16_t* dfine, int f_ld, int f_off, Vol vf,
                              uint16_t* dcoarse, int c_ld, int c_off, Vol vc, int batch, int c);
// mean squared error over the first `c` of 16 fp32 channels; writes
// dpred = 2 (pred - target) / count in bf16 (zero in the padding channels)
// and the loss (sum of squares / count) into *loss_dev
hipError_t launch_mse(hipStream_t s, const float* pred, const float* target, int64_t voxels, int c,
                      uint16_t* dpred, float* part, float* loss_dev);
// Adam with bias correction folded into lr_t
hipError_t launch_adam(hipStream_t s, float* p, const float* g, float* m, float* v, int64_t n,
                       float lr_t, float beta1, float beta2, float eps);
// x[3i + c] = scale * (pred[i][c] + base_w * feat[i][base_ch + c]) (c < 3;
// no base term when base_ch < 0), zero where feature channel mask_ch of
// node i exceeds 1/2 (the clamped nodes); one sample
hipError_t launch_to_dofs(hipStream_t s, const float* pred, const uint16_t* feat, int mask_ch,
                          int base_ch, double base_w, int64_t nodes, double scale, double* x);
// zero-fill of a device buffer (a kernel, so it may sit inside a captured graph)
hipError_t launch_fill0(hipStream_t s, void* p, int64_t bytes);

}  // namespace nn
}  // namespace gf
