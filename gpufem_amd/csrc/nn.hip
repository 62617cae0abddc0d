// Layer kernels of the learned warm start (BF16 storage, FP32 accumulate).
//
// The convolutions are implicit GEMMs on the gfx950 matrix cores with
// v_mfma_f32_16x16x16_bf16: one instruction multiplies a 16x16 A tile by a
// 16x16 B tile.  Lane l holds A[l%16][4(l/16)+e] and B[4(l/16)+e][l%16]
// (e = 0..3) and receives D[4(l/16)+e][l%16].  With A = weights (output
// channel x input channel) and B = gathered activations (input channel x
// voxel), each lane ends up owning four consecutive output channels of one
// voxel, so the epilogue stores 8 contiguous bytes per lane, and the B
// operand is one 8-byte load of four consecutive channels per lane.
#include <cfloat>

#include "nn.hpp"

namespace gf {
namespace nn {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kNT = 256;

__device__ inline float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// round to nearest even; NaN stays NaN
__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

__device__ inline f4 mfma16(s4 a, s4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// source coordinate along one axis for output coordinate v and tap offset t
__device__ inline int src_coord(int v, int t, int s, int transposed, int n) {
  int c;
  if (!transposed) {
    c = s * v + t;
  } else {
    const int d = v - t;
    if (d < 0 || (d % s) != 0) return -1;
    c = d / s;
  }
  return (c >= 0 && c < n) ? c : -1;
}

// ---- convolution ----------------------------------------------------------------
// A wave owns VB blocks of 16 voxels and CT tiles of 16 output channels; a
// workgroup is four such waves on consecutive voxels.
template <int VB, int CT>
__global__ __launch_bounds__(kNT) void k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t nout = a.vout.nodes(), nin = a.vin.nodes();
  const int64_t total = int64_t(a.batch) * nout;
  const int ncb = a.cin >> 4, nct = a.cout >> 4;
  const int ct0 = blockIdx.y * CT;
  const int kq = (lane >> 4) << 2;
  int64_t vox[VB], ibase[VB];
  int vz[VB], vy[VB], vx[VB];
  bool ok[VB];
#pragma unroll
  for (int j = 0; j < VB; ++j) {
    const int64_t v = (int64_t(blockIdx.x) * 4 + wave) * (16 * VB) + j * 16 + (lane & 15);
    vox[j] = v;
    ok[j] = v < total;
    const int64_t vv = ok[j] ? v : 0;
    const int64_t b = vv / nout;
    int64_t r = vv - b * nout;
    vx[j] = int(r % a.vout.nx);
    r /= a.vout.nx;
    vy[j] = int(r % a.vout.ny);
    vz[j] = int(r / a.vout.ny);
    ibase[j] = b * nin;
  }
  f4 acc[VB][CT];
#pragma unroll
  for (int j = 0; j < VB; ++j)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[j][c] = f4{0.f, 0.f, 0.f, 0.f};
  const s4 zero4 = {0, 0, 0, 0};
  for (int tap = 0; tap < 27; ++tap) {
    const int tz = tap / 9 - 1, ty = (tap / 3) % 3 - 1, tx = tap % 3 - 1;
    int64_t src[VB];
#pragma unroll
    for (int j = 0; j < VB; ++j) {
      const int sz = src_coord(vz[j], tz, a.stride, a.transposed, a.vin.nz);
      const int sy = src_coord(vy[j], ty, a.stride, a.transposed, a.vin.ny);
      const int sx = src_coord(vx[j], tx, a.stride, a.transposed, a.vin.nx);
      src[j] = (ok[j] && sz >= 0 && sy >= 0 && sx >= 0)
                   ? (ibase[j] + (int64_t(sz) * a.vin.ny + sy) * a.vin.nx + sx) * a.in_ld + a.in_off + kq
                   : -1;
    }
    const uint16_t* wp = a.wf + ((int64_t(tap * ncb) * nct + ct0) * 64 + lane) * 4;
    for (int cb = 0; cb < ncb; ++cb) {
      s4 bfr[VB];
#pragma unroll
      for (int j = 0; j < VB; ++j)
        bfr[j] = src[j] >= 0 ? *reinterpret_cast<const s4*>(a.in + src[j] + cb * 16) : zero4;
      const uint16_t* wc = wp + int64_t(cb) * nct * 256;
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        const s4 af = *reinterpret_cast<const s4*>(wc + c * 256);
#pragma unroll
        for (int j = 0; j < VB; ++j) acc[j][c] = mfma16(af, bfr[j], acc[j][c]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < VB; ++j) {
    if (!ok[j]) continue;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int co = (ct0 + c) * 16 + kq;
      float r[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        float t = acc[j][c][e] + (a.bias ? a.bias[co + e] : 0.f);
        if (a.act) t = t > 0.f ? t : a.slope * t;
        r[e] = t;
      }
      if (a.out32) {
        float4* o = reinterpret_cast<float4*>(a.out32 + vox[j] * a.out_ld + a.out_off + co);
        float4 w = make_float4(r[0], r[1], r[2], r[3]);
        if (a.accumulate) {
          const float4 old = *o;
          w.x += old.x;
          w.y += old.y;
          w.z += old.z;
          w.w += old.w;
        }
        *o = w;
      } else {
        uint2* o = reinterpret_cast<uint2*>(a.out + vox[j] * a.out_ld + a.out_off + co);
        if (a.accumulate) {
          const uint2 old = *o;
          r[0] += bf2f(uint16_t(old.x & 0xffffu));
          r[1] += bf2f(uint16_t(old.x >> 16));
          r[2] += bf2f(uint16_t(old.y & 0xffffu));
          r[3] += bf2f(uint16_t(old.y >> 16));
        }
        uint2 w;
        w.x = uint32_t(f2bf(r[0])) | (uint32_t(f2bf(r[1])) << 16);
        w.y = uint32_t(f2bf(r[2])) | (uint32_t(f2bf(r[3])) << 16);
        *o = w;
      }
    }
  }
}

// ---- weight gradient ---------------------------------------------------------------
// Workgroup (chunk, tap): stages 64-voxel tiles of the gathered input and of
// dz in LDS transposed ([channel][voxel], padded rows), so that each MFMA
// operand (four consecutive voxels of one channel) is one 8-byte LDS read;
// every wave then multiplies its 16x16 (ci, co) tiles with the voxel axis as
// the reduction axis.
constexpr int kWgTile = 64;
constexpr int kWgRow = kWgTile + 4;  // LDS row stride (elements)

template <int TPW>
__global__ __launch_bounds__(kNT) void k_wgrad(WgradArgs a) {
  extern __shared__ uint16_t sm[];
  uint16_t* sx = sm;                       // [cin][kWgRow]
  uint16_t* sd = sm + a.cin * kWgRow;      // [cout][kWgRow]
  const int tap = blockIdx.y;
  const int tz = tap / 9 - 1, ty = (tap / 3) % 3 - 1, tx = tap % 3 - 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nco = a.cout >> 4, ntile = (a.cin >> 4) * nco;
  const int64_t nout = a.vout.nodes(), nin = a.vin.nodes();
  const int64_t total = int64_t(a.batch) * nout;
  const int64_t v0 = int64_t(blockIdx.x) * a.chunk;
  const int64_t v1 = v0 + a.chunk < total ? v0 + a.chunk : total;
  f4 acc[TPW];
#pragma unroll
  for (int j = 0; j < TPW; ++j) acc[j] = f4{0.f, 0.f, 0.f, 0.f};
  const int cq_in = a.cin >> 2, cq_d = a.cout >> 2;
  for (int64_t base = v0; base < v1; base += kWgTile) {
    __syncthreads();
    for (int i = threadIdx.x; i < kWgTile * cq_in; i += kNT) {
      const int r = i % kWgTile, q = i / kWgTile;  // consecutive threads: consecutive voxels
      const int64_t v = base + r;
      uint2 val = make_uint2(0u, 0u);
      if (v < v1) {
        const int64_t b = v / nout;
        int64_t rr = v - b * nout;
        const int x = int(rr % a.vout.nx);
        rr /= a.vout.nx;
        const int y = int(rr % a.vout.ny), z = int(rr / a.vout.ny);
        const int sz = src_coord(z, tz, a.stride, 0, a.vin.nz);
        const int sy = src_coord(y, ty, a.stride, 0, a.vin.ny);
        const int sx2 = src_coord(x, tx, a.stride, 0, a.vin.nx);
        if (sz >= 0 && sy >= 0 && sx2 >= 0)
          val = *reinterpret_cast<const uint2*>(
              a.x + (b * nin + (int64_t(sz) * a.vin.ny + sy) * a.vin.nx + sx2) * a.x_ld + a.x_off + q * 4);
      }
      uint16_t* col = sx + (q * 4) * kWgRow + r;
      col[0] = uint16_t(val.x & 0xffffu);
      col[kWgRow] = uint16_t(val.x >> 16);
      col[2 * kWgRow] = uint16_t(val.y & 0xffffu);
      col[3 * kWgRow] = uint16_t(val.y >> 16);
    }
    for (int i = threadIdx.x; i < kWgTile * cq_d; i += kNT) {
      const int r = i % kWgTile, q = i / kWgTile;
      const int64_t v = base + r;
      uint2 val = make_uint2(0u, 0u);
      if (v < v1) val = *reinterpret_cast<const uint2*>(a.dz + v * a.dz_ld + a.dz_off + q * 4);
      uint16_t* col = sd + (q * 4) * kWgRow + r;
      col[0] = uint16_t(val.x & 0xffffu);
      col[kWgRow] = uint16_t(val.x >> 16);
      col[2 * kWgRow] = uint16_t(val.y & 0xffffu);
      col[3 * kWgRow] = uint16_t(val.y >> 16);
    }
    __syncthreads();
#pragma unroll 1
    for (int kb = 0; kb < kWgTile / 16; ++kb) {
      const int c0 = kb * 16 + ((lane >> 4) << 2);
#pragma unroll
      for (int j = 0; j < TPW; ++j) {
        const int t = wave + 4 * j;
        if (t < ntile) {
          const int it = t / nco, ot = t - it * nco;
          const s4 A = *reinterpret_cast<const s4*>(sx + (it * 16 + (lane & 15)) * kWgRow + c0);
          const s4 B = *reinterpret_cast<const s4*>(sd + (ot * 16 + (lane & 15)) * kWgRow + c0);
          acc[j] = mfma16(A, B, acc[j]);
        }
      }
    }
  }
  float* P = a.part + (int64_t(blockIdx.x) * 27 + tap) * a.cin * a.cout;
#pragma unroll
  for (int j = 0; j < TPW; ++j) {
    const int t = wave + 4 * j;
    if (t < ntile) {
      const int it = t / nco, ot = t - it * nco;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int ci = it * 16 + ((lane >> 4) << 2) + e, co = ot * 16 + (lane & 15);
        P[ci * a.cout + co] = acc[j][e];
      }
    }
  }
}

// out[i] = sum_c part[c * len + i] (fixed order)
__global__ void k_reduce_parts(const float* part, int np, int64_t len, float* out) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len;
       i += int64_t(gridDim.x) * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < np; ++c) s += part[c * len + i];
    out[i] = s;
  }
}

// ---- weights -> MFMA A fragments -------------------------------------------------
__global__ void k_arrange(const float* w, uint16_t* wf, int cin, int cout, int transposed) {
  const int Cc = transposed ? cout : cin, Co = transposed ? cin : cout;
  const int64_t total = int64_t(27) * Cc * Co;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int e = int(idx & 3), lane = int((idx >> 2) & 63);
    int64_t rest = idx >> 8;
    const int ot = int(rest % (Co >> 4));
    rest /= (Co >> 4);
    const int cb = int(rest % (Cc >> 4));
    const int tap = int(rest / (Cc >> 4));
    const int c = cb * 16 + ((lane >> 4) << 2) + e, o = ot * 16 + (lane & 15);
    const float v = transposed ? w[(int64_t(tap) * cin + o) * cout + c]
                               : w[(int64_t(tap) * cin + c) * cout + o];
    wf[idx] = f2bf(v);
  }
}

// ---- element-wise and small reductions -----------------------------------------
__global__ void k_act_bwd(uint16_t* dy, int dy_ld, int dy_off, const uint16_t* y, int y_ld,
                          int y_off, int64_t voxels, int cq, float slope) {
  const int64_t total = voxels * cq;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = i / cq;
    const int q = int(i - v * cq);
    uint2* pd = reinterpret_cast<uint2*>(dy + v * dy_ld + dy_off + q * 4);
    const uint2 yy = *reinterpret_cast<const uint2*>(y + v * y_ld + y_off + q * 4);
    uint2 d = *pd;
    const uint16_t ys[4] = {uint16_t(yy.x & 0xffffu), uint16_t(yy.x >> 16), uint16_t(yy.y & 0xffffu),
                            uint16_t(yy.y >> 16)};
    uint16_t ds[4] = {uint16_t(d.x & 0xffffu), uint16_t(d.x >> 16), uint16_t(d.y & 0xffffu),
                      uint16_t(d.y >> 16)};
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (!(bf2f(ys[e]) > 0.f)) ds[e] = f2bf(slope * bf2f(ds[e]));
    d.x = uint32_t(ds[0]) | (uint32_t(ds[1]) << 16);
    d.y = uint32_t(ds[2]) | (uint32_t(ds[3]) << 16);
    *pd = d;
  }
}

// per-block channel sums; c divides kNT
__global__ __launch_bounds__(kNT) void k_bias_part(const uint16_t* dz, int ld, int off,
                                                   int64_t voxels, int c, float* part) {
  __shared__ float sh[kNT];
  const int ch = threadIdx.x % c, vp = threadIdx.x / c, nvp = kNT / c;
  float s = 0.f;
  for (int64_t v = int64_t(blockIdx.x) * nvp + vp; v < voxels; v += int64_t(gridDim.x) * nvp)
    s += bf2f(dz[v * ld + off + ch]);
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < c) {
    float t = 0.f;
    for (int k = 0; k < nvp; ++k) t += sh[k * c + threadIdx.x];
    part[int64_t(blockIdx.x) * c + threadIdx.x] = t;
  }
}

__device__ inline void axis_up(int f, int nc, int* i0, int* i1) {
  // fine f -> coarse neighbours (weight 1 if f even, 1/2 each if odd)
  if ((f & 1) == 0) {
    *i0 = f >> 1;
    *i1 = -1;
  } else {
    *i0 = (f - 1) >> 1;
    *i1 = (f + 1) >> 1;
    if (*i1 >= nc) *i1 = -1;
  }
}

__global__ void k_prolong(const uint16_t* in, int in_ld, int in_off, Vol vc, uint16_t* out,
                          int out_ld, int out_off, Vol vf, int batch, int cq) {
  const int64_t nf = vf.nodes(), total = int64_t(batch) * nf * cq;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = i / cq;
    const int q = int(i - v * cq);
    const int64_t b = v / nf;
    int64_t r = v - b * nf;
    const int x = int(r % vf.nx);
    r /= vf.nx;
    const int y = int(r % vf.ny), z = int(r / vf.ny);
    int cz[2], cy[2], cx[2];
    axis_up(z, vc.nz, &cz[0], &cz[1]);
    axis_up(y, vc.ny, &cy[0], &cy[1]);
    axis_up(x, vc.nx, &cx[0], &cx[1]);
    const float wz = cz[1] >= 0 ? 0.5f : 1.f, wy = cy[1] >= 0 ? 0.5f : 1.f,
                wx = cx[1] >= 0 ? 0.5f : 1.f;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int a = 0; a < 2; ++a) {
      if (cz[a] < 0) continue;
      for (int bb = 0; bb < 2; ++bb) {
        if (cy[bb] < 0) continue;
        for (int c = 0; c < 2; ++c) {
          if (cx[c] < 0) continue;
          const float w = wz * wy * wx;
          const uint2 u = *reinterpret_cast<const uint2*>(
              in + (b * vc.nodes() + (int64_t(cz[a]) * vc.ny + cy[bb]) * vc.nx + cx[c]) * in_ld +
              in_off + q * 4);
          acc[0] += w * bf2f(uint16_t(u.x & 0xffffu));
          acc[1] += w * bf2f(uint16_t(u.x >> 16));
          acc[2] += w * bf2f(uint16_t(u.y & 0xffffu));
          acc[3] += w * bf2f(uint16_t(u.y >> 16));
        }
      }
    }
    uint2 o;
    o.x = uint32_t(f2bf(acc[0])) | (uint32_t(f2bf(acc[1])) << 16);
    o.y = uint32_t(f2bf(acc[2])) | (uint32_t(f2bf(acc[3])) << 16);
    *reinterpret_cast<uint2*>(out + v * out_ld + out_off + q * 4) = o;
  }
}

__global__ void k_prolong_bwd(const uint16_t* df, int f_ld, int f_off, Vol vf, uint16_t* dc,
                              int c_ld, int c_off, Vol vc, int batch, int cq) {
  const int64_t nc = vc.nodes(), total = int64_t(batch) * nc * cq;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t v = i / cq;
    const int q = int(i - v * cq);
    const int64_t b = v / nc;
    int64_t r = v - b * nc;
    const int X = int(r % vc.nx);
    r /= vc.nx;
    const int Y = int(r % vc.ny), Z = int(r / vc.ny);
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int dz = -1; dz <= 1; ++dz) {
      const int fz = 2 * Z + dz;
      if (fz < 0 || fz >= vf.nz) continue;
      const float wz = dz == 0 ? 1.f : 0.5f;
      for (int dy = -1; dy <= 1; ++dy) {
        const int fy = 2 * Y + dy;
        if (fy < 0 || fy >= vf.ny) continue;
        const float wy = dy == 0 ? 1.f : 0.5f;
        for (int dx = -1; dx <= 1; ++dx) {
          const int fx = 2 * X + dx;
          if (fx < 0 || fx >= vf.nx) continue;
          const float w = wz * wy * (dx == 0 ? 1.f : 0.5f);
          const uint2 u = *reinterpret_cast<const uint2*>(
              df + (b * vf.nodes() + (int64_t(fz) * vf.ny + fy) * vf.nx + fx) * f_ld + f_off + q * 4);
          acc[0] += w * bf2f(uint16_t(u.x & 0xffffu));
          acc[1] += w * bf2f(uint16_t(u.x >> 16));
          acc[2] += w * bf2f(uint16_t(u.y & 0xffffu));
          acc[3] += w * bf2f(uint16_t(u.y >> 16));
        }
      }
    }
    uint2 o;
    o.x = uint32_t(f2bf(acc[0])) | (uint32_t(f2bf(acc[1])) << 16);
    o.y = uint32_t(f2bf(acc[2])) | (uint32_t(f2bf(acc[3])) << 16);
    *reinterpret_cast<uint2*>(dc + v * c_ld + c_off + q * 4) = o;
  }
}

__global__ __launch_bounds__(kNT) void k_mse(const float* pred, const float* target,
                                             int64_t voxels, int c, uint16_t* dpred,
                                             float* part) {
  __shared__ float sh[kNT];
  const float scale = 2.f / float(voxels * c);
  float s = 0.f;
  for (int64_t v = int64_t(blockIdx.x) * kNT + threadIdx.x; v < voxels;
       v += int64_t(gridDim.x) * kNT) {
    uint16_t g[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) g[e] = 0;
    for (int e = 0; e < c; ++e) {
      const float d = pred[v * 16 + e] - target[v * 4 + e];
      s += d * d;
      g[e] = f2bf(scale * d);
    }
    uint4* o = reinterpret_cast<uint4*>(dpred + v * 16);
    o[0] = make_uint4(g[0] | (uint32_t(g[1]) << 16), g[2] | (uint32_t(g[3]) << 16),
                      g[4] | (uint32_t(g[5]) << 16), g[6] | (uint32_t(g[7]) << 16));
    o[1] = make_uint4(g[8] | (uint32_t(g[9]) << 16), g[10] | (uint32_t(g[11]) << 16),
                      g[12] | (uint32_t(g[13]) << 16), g[14] | (uint32_t(g[15]) << 16));
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = kNT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_mse_finish(const float* part, int np, float inv_count, float* loss) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < np; ++i) s += part[i];
    *loss = s * inv_count;
  }
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr_t,
                       float b1, float b2, float eps) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps);
  }
}

__global__ void k_to_dofs(const float* pred, const uint16_t* feat, int mask_ch, int base_ch,
                          double base_w, int64_t nodes, double scale, double* x) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < nodes;
       i += int64_t(gridDim.x) * blockDim.x) {
    const bool fixed = bf2f(feat[i * 16 + mask_ch]) > 0.5f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = double(pred[i * 16 + c]);
      if (base_ch >= 0) v += base_w * double(bf2f(feat[i * 16 + base_ch + c]));
      x[i * 3 + c] = fixed ? 0.0 : scale * v;
    }
  }
}

__global__ void k_fill0(uint32_t* p, int64_t words) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < words;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = 0u;
}

inline int grid_for(int64_t work, int cap = 8192) {
  int64_t b = (work + kNT - 1) / kNT;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return int(b);
}

template <int VB>
hipError_t conv_vb(hipStream_t s, const ConvArgs& a) {
  const int nct = a.cout >> 4;
  const int64_t total = int64_t(a.batch) * a.vout.nodes();
  const int64_t gx = (total + 64 * VB - 1) / (64 * VB);
  if (gx > 0x7fffffff) return hipErrorInvalidValue;
  if (nct % 4 == 0) {
    hipLaunchKernelGGL((k_conv<VB, 4>), dim3(unsigned(gx), nct / 4), dim3(kNT), 0, s, a);
  } else if (nct % 3 == 0) {
    hipLaunchKernelGGL((k_conv<VB, 3>), dim3(unsigned(gx), nct / 3), dim3(kNT), 0, s, a);
  } else if (nct % 2 == 0) {
    hipLaunchKernelGGL((k_conv<VB, 2>), dim3(unsigned(gx), nct / 2), dim3(kNT), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_conv<VB, 1>), dim3(unsigned(gx), nct), dim3(kNT), 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_conv(hipStream_t s, const ConvArgs& a) {
  if (a.cin < 16 || a.cout < 16 || (a.cin & 15) || (a.cout & 15) || (a.in_ld & 3) ||
      (a.in_off & 3) || (a.out_ld & 3) || (a.out_off & 3) || a.batch < 1 ||
      (a.stride != 1 && a.stride != 2) || (!a.out) == (!a.out32) ||
      a.in_off + a.cin > a.in_ld || a.out_off + a.cout > a.out_ld)
    return hipErrorInvalidValue;
  return conv_vb<2>(s, a);
}

int wgrad_chunks(int64_t voxels) {
  int64_t c = (voxels + 2047) / 2048;
  if (c < 1) c = 1;
  if (c > 64) c = 64;
  return int(c);
}

hipError_t launch_wgrad(hipStream_t s, const WgradArgs& a, float* dw) {
  if ((a.cin & 15) || (a.cout & 15) || a.cin < 16 || a.cout < 16 || (a.chunk % kWgTile) ||
      a.nchunk < 1 || (a.x_ld & 3) || (a.x_off & 3) || (a.dz_ld & 3) || (a.dz_off & 3))
    return hipErrorInvalidValue;
  const int64_t total = int64_t(a.batch) * a.vout.nodes();
  if (int64_t(a.nchunk) * a.chunk < total) return hipErrorInvalidValue;
  const int ntile = (a.cin >> 4) * (a.cout >> 4);
  const int tpw = (ntile + 3) / 4;
  const size_t lds = size_t(kWgRow) * (a.cin + a.cout) * sizeof(uint16_t);
  if (lds > 64 * 1024) return hipErrorInvalidValue;
  const dim3 g(a.nchunk, 27), t(kNT);
  if (tpw <= 1) hipLaunchKernelGGL(k_wgrad<1>, g, t, lds, s, a);
  else if (tpw <= 2) hipLaunchKernelGGL(k_wgrad<2>, g, t, lds, s, a);
  else if (tpw <= 4) hipLaunchKernelGGL(k_wgrad<4>, g, t, lds, s, a);
  else if (tpw <= 8) hipLaunchKernelGGL(k_wgrad<8>, g, t, lds, s, a);
  else if (tpw <= 12) hipLaunchKernelGGL(k_wgrad<12>, g, t, lds, s, a);
  else return hipErrorInvalidValue;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t len = int64_t(27) * a.cin * a.cout;
  hipLaunchKernelGGL(k_reduce_parts, dim3(grid_for(len)), dim3(kNT), 0, s, a.part, a.nchunk, len, dw);
  return hipGetLastError();
}

hipError_t launch_arrange(hipStream_t s, const float* w, uint16_t* wf, int cin, int cout,
                          int transposed) {
  if ((cin & 15) || (cout & 15)) return hipErrorInvalidValue;
  const int64_t total = int64_t(27) * cin * cout;
  hipLaunchKernelGGL(k_arrange, dim3(grid_for(total)), dim3(kNT), 0, s, w, wf, cin, cout, transposed);
  return hipGetLastError();
}

hipError_t launch_act_bwd(hipStream_t s, uint16_t* dy, int dy_ld, int dy_off, const uint16_t* y,
                          int y_ld, int y_off, int64_t voxels, int c, float slope) {
  if ((c & 3) || (dy_ld & 3) || (dy_off & 3) || (y_ld & 3) || (y_off & 3)) return hipErrorInvalidValue;
  const int cq = c >> 2;
  hipLaunchKernelGGL(k_act_bwd, dim3(grid_for(voxels * cq)), dim3(kNT), 0, s, dy, dy_ld, dy_off, y,
                     y_ld, y_off, voxels, cq, slope);
  return hipGetLastError();
}

hipError_t launch_bias_grad(hipStream_t s, const uint16_t* dz, int ld, int off, int64_t voxels,
                            int c, float* part, float* db) {
  if (c < 1 || c > kNT || (kNT % c)) return hipErrorInvalidValue;
  const int nb = 256;
  hipLaunchKernelGGL(k_bias_part, dim3(nb), dim3(kNT), 0, s, dz, ld, off, voxels, c, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_reduce_parts, dim3(1), dim3(kNT), 0, s, part, nb, int64_t(c), db);
  return hipGetLastError();
}

hipError_t launch_prolong(hipStream_t s, const uint16_t* in, int in_ld, int in_off, Vol vc,
                          uint16_t* out, int out_ld, int out_off, Vol vf, int batch, int c) {
  if ((c & 3) || vf.nx != 2 * vc.nx - 1 || vf.ny != 2 * vc.ny - 1 || vf.nz != 2 * vc.nz - 1)
    return hipErrorInvalidValue;
  const int cq = c >> 2;
  hipLaunchKernelGGL(k_prolong, dim3(grid_for(int64_t(batch) * vf.nodes() * cq)), dim3(kNT), 0, s,
                     in, in_ld, in_off, vc, out, out_ld, out_off, vf, batch, cq);
  return hipGetLastError();
}

hipError_t launch_prolong_bwd(hipStream_t s, const uint16_t* dfine, int f_ld, int f_off, Vol vf,
                              uint16_t* dcoarse, int c_ld, int c_off, Vol vc, int batch, int c) {
  if ((c & 3) || vf.nx != 2 * vc.nx - 1 || vf.ny != 2 * vc.ny - 1 || vf.nz != 2 * vc.nz - 1)
    return hipErrorInvalidValue;
  const int cq = c >> 2;
  hipLaunchKernelGGL(k_prolong_bwd, dim3(grid_for(int64_t(batch) * vc.nodes() * cq)), dim3(kNT), 0,
                     s, dfine, f_ld, f_off, vf, dcoarse, c_ld, c_off, vc, batch, cq);
  return hipGetLastError();
}

hipError_t launch_mse(hipStream_t s, const float* pred, const float* target, int64_t voxels, int c,
                      uint16_t* dpred, float* part, float* loss_dev) {
  if (c < 1 || c > 4) return hipErrorInvalidValue;
  const int nb = grid_for(voxels, 1024);
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(kNT), 0, s, pred, target, voxels, c, dpred, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mse_finish, dim3(1), dim3(64), 0, s, part, nb, 1.f / float(voxels * c), loss_dev);
  return hipGetLastError();
}

hipError_t launch_adam(hipStream_t s, float* p, const float* g, float* m, float* v, int64_t n,
                       float lr_t, float beta1, float beta2, float eps) {
  hipLaunchKernelGGL(k_adam, dim3(grid_for(n)), dim3(kNT), 0, s, p, g, m, v, n, lr_t, beta1, beta2, eps);
  return hipGetLastError();
}

hipError_t launch_to_dofs(hipStream_t s, const float* pred, const uint16_t* feat, int mask_ch,
                          int base_ch, double base_w, int64_t nodes, double scale, double* x) {
  if (mask_ch < 0 || mask_ch >= 16 || base_ch < -1 || base_ch > 13) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_to_dofs, dim3(grid_for(nodes)), dim3(kNT), 0, s, pred, feat, mask_ch,
                     base_ch, base_w, nodes, scale, x);
  return hipGetLastError();
}

hipError_t launch_fill0(hipStream_t s, void* p, int64_t bytes) {
  if (bytes & 3) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fill0, dim3(grid_for(bytes / 4)), dim3(kNT), 0, s, static_cast<uint32_t*>(p), bytes / 4);
  return hipGetLastError();
}

}  // namespace nn
}  // namespace gf
