// x0net_kernels.hip -- kernels of the initial-guess network (3D U-Net).
//
// Tensors are channel-last [Z][Y][X][C] with C a multiple of 16; network
// activations and gradients are BF16 (raw 16-bit storage), accumulation is
// FP32.  The 3x3x3 convolutions are implicit GEMMs on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16):
//   forward       y[v][co] = sum_{tap,ci} x[v+tap-1][ci] w[co][tap*Cin+ci]
//                 M = voxels, N = Cout, K = 27*Cin (padded to 32)
//   grad of input dx = the forward kernel applied to dy with the flipped,
//                 transposed weights w'[ci][tap'*Cout+co] = w[co][(26-tap')*Cin+ci]
//   grad of weight dw[co][tap*Cin+ci] = sum_v dy[v][co] x[v+tap-1][ci]
//                 M = Cout, N = Cin per tap, K = voxels, partial sums per
//                 slab of z-slices, then a fixed-order reduction
// Operand layout of the 16x16x32 instruction (one wave): lane l holds
// A[l%16][8(l/16)+j] and B[8(l/16)+j][l%16] (j < 8) and receives
// D[4(l/16)+i][l%16] (i < 4).
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <algorithm>
#include <cmath>

#include "emmi_mgfd.h"
#include "timing.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

__device__ __forceinline__ float bf2f(bf16 v) { return __uint_as_float(((uint32_t)v) << 16); }
__device__ __forceinline__ bf16 f2bf(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

__device__ __forceinline__ bf16x8 load8(const bf16* p) {
  const uint4 v = *reinterpret_cast<const uint4*>(p);
  return __builtin_bit_cast(bf16x8, v);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0)); }

enum Epilogue { EPI_LEAKY = 0, EPI_LINEAR = 1, EPI_DACT = 2, EPI_LINEAR_BF16 = 3 };

// One wave: 16*MT consecutive voxels along x of row (z, y) and all 16*NT
// output channels.  Block: 4 waves on rows y0..y0+3.
template <int MT, int NT>
__global__ void __launch_bounds__(256) k_conv_fwd(const bf16* __restrict__ x, int Z, int Y, int X, int Cin,
                                                  const bf16* __restrict__ w, int Kpad,
                                                  const float* __restrict__ bias, int epi, float slope,
                                                  const bf16* __restrict__ act, bf16* __restrict__ out_bf,
                                                  float* __restrict__ out_f, int Cout_store) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int x0 = blockIdx.x * 16 * MT;
  const int y = blockIdx.y * 4 + wave;
  const int z = blockIdx.z;
  if (y >= Y) return;
  f32x4 acc[MT][NT];
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int n = 0; n < NT; n++) acc[m][n] = f32x4{0, 0, 0, 0};
  const int K27 = 27 * Cin;
  const int row = lane & 15, kq = (lane >> 4) * 8;
  for (int k0 = 0; k0 < Kpad; k0 += 32) {
    const int k = k0 + kq;
    bf16x8 a[MT];
    if (k < K27) {
      const int tap = k / Cin, ci = k - tap * Cin;
      const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
      const int zz = z + dz, yy = y + dy;
      const bool zy_ok = zz >= 0 && zz < Z && yy >= 0 && yy < Y;
#pragma unroll
      for (int m = 0; m < MT; m++) {
        const int xx = x0 + 16 * m + row + dx;
        if (zy_ok && xx >= 0 && xx < X)
          a[m] = load8(x + (((long)zz * Y + yy) * X + xx) * Cin + ci);
        else
          a[m] = zero8();
      }
    } else {
#pragma unroll
      for (int m = 0; m < MT; m++) a[m] = zero8();
    }
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const bf16x8 b = load8(w + (long)(16 * n + row) * Kpad + k);
#pragma unroll
      for (int m = 0; m < MT; m++) acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], b, acc[m][n], 0, 0, 0);
    }
  }
  // epilogue: lane holds voxels x0+16m+4(l/16)+i, channel 16n+(l%16)
  const int Cout = 16 * NT;
#pragma unroll
  for (int m = 0; m < MT; m++)
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const int co = 16 * n + row;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int xx = x0 + 16 * m + 4 * (lane >> 4) + i;
        if (xx >= X) continue;
        const long v = ((long)z * Y + y) * X + xx;
        float r = acc[m][n][i];
        if (epi == EPI_LEAKY) {
          r += bias[co];
          r = r > 0.f ? r : slope * r;
          out_bf[v * Cout + co] = f2bf(r);
        } else if (epi == EPI_LINEAR) {
          r += bias ? bias[co] : 0.f;
          if (co < Cout_store) out_f[v * Cout_store + co] = r;
        } else if (epi == EPI_LINEAR_BF16) {
          out_bf[v * Cout + co] = f2bf(r);
        } else {
          // gradient through the activation of the tensor this one feeds
          const float av = bf2f(act[v * Cout + co]);
          out_bf[v * Cout + co] = f2bf(av > 0.f ? r : slope * r);
        }
      }
    }
}

// dw partials: block = 4 waves, one (tap, 16-channel input tile, 16-channel
// output tile) per block, the voxels of slab blockIdx.y in tiles of 128.
// A tile of dy (16 output channels) and of x at the tap's offset (16 input
// channels, zero outside the grid) is loaded with 16-byte loads (a thread: 8
// channels of one voxel) and stored channel-major in LDS, so that a lane's
// 8 consecutive voxels of one channel -- its share of an MFMA operand -- are
// one aligned 16-byte LDS read.  Rows are padded to 136 values so the 16
// channel rows start in different banks.  The next tile's loads are in
// flight while the current one is multiplied.  Wave w takes voxels
// [32 w, 32 w + 32) of each tile.
constexpr int DW_TV = 128, DW_ROW = DW_TV + 8;
__global__ void __launch_bounds__(256) k_conv_dw(const bf16* __restrict__ x, const bf16* __restrict__ dy, int Z, int Y,
                                                 int X, int Cin, int Cout, int zslab, float* __restrict__ part) {
  __shared__ __align__(16) bf16 As[16 * DW_ROW];
  __shared__ __align__(16) bf16 Bs[16 * DW_ROW];
  __shared__ float red[4][64][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nci = Cin / 16, nco = Cout / 16;
  const int t = blockIdx.x;
  const int tap = t / (nci * nco);
  const int ci0 = ((t / nco) % nci) * 16, co0 = (t % nco) * 16;
  const int dz = tap / 9 - 1, dyy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int z0 = blockIdx.y * zslab, z1 = min(Z, z0 + zslab);
  // voxels of a slab fit 32-bit indices (the launcher checks the volume)
  const int nv = (z1 - z0) * Y * X;
  const int base = z0 * Y * X;   // flat index of the slab's first voxel
  // staging role: voxel vi of the tile, channels 8 h .. 8 h + 7
  const int vi = threadIdx.x >> 1, h = threadIdx.x & 1;
  auto load = [&](int v0, uint4& ra, uint4& rb) {
    ra = make_uint4(0, 0, 0, 0);
    rb = make_uint4(0, 0, 0, 0);
    const int v = v0 + vi;
    if (v < nv) {
      const int g = base + v;
      ra = *reinterpret_cast<const uint4*>(dy + (long)g * Cout + co0 + 8 * h);
      const int xx = g % X, q = g / X, yy = q % Y, zz = q / Y;
      const int sz = zz + dz, sy = yy + dyy, sx = xx + dx;
      if (sz >= 0 && sz < Z && sy >= 0 && sy < Y && sx >= 0 && sx < X)
        rb = *reinterpret_cast<const uint4*>(x + ((long)(sz * Y + sy) * X + sx) * Cin + ci0 + 8 * h);
    }
  };
  auto put = [&](const uint4& ra, const uint4& rb) {
    const bf16* pa = reinterpret_cast<const bf16*>(&ra);
    const bf16* pb = reinterpret_cast<const bf16*>(&rb);
#pragma unroll
    for (int c = 0; c < 8; c++) {
      As[(8 * h + c) * DW_ROW + vi] = pa[c];
      Bs[(8 * h + c) * DW_ROW + vi] = pb[c];
    }
  };
  const int col = lane & 15, kq = (lane >> 4) * 8;
  f32x4 acc = {0, 0, 0, 0};
  uint4 ra, rb;
  load(0, ra, rb);
  for (int v0 = 0; v0 < nv; v0 += DW_TV) {
    __syncthreads();   // the previous tile's operands are read
    put(ra, rb);
    __syncthreads();
    if (v0 + DW_TV < nv) load(v0 + DW_TV, ra, rb);
    // A = dy^T (rows co, cols voxels), B = x_tap (rows voxels, cols ci)
    const int o = col * DW_ROW + 32 * wave + kq;
    const bf16x8 a = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&As[o]));
    const bf16x8 b = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(&Bs[o]));
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) red[wave][lane][i] = acc[i];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float s = red[0][lane][i] + red[1][lane][i] + red[2][lane][i] + red[3][lane][i];
      const int co = co0 + 4 * (lane >> 4) + i, ci = ci0 + col;
      part[((long)blockIdx.y * Cout + co) * (27L * Cin) + (long)tap * Cin + ci] = s;
    }
  }
}

__global__ void k_reduce_slabs(const float* __restrict__ part, int nslab, long n, float* __restrict__ dw) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int q = 0; q < nslab; q++) s += part[q * n + i];
  dw[i] = s;
}

// bias gradient db[co] = sum over voxels of dy[v][co] (channels fastest), in
// two passes of fixed order: block b sums the voxels [b chunk, (b+1) chunk)
// of every channel, 256 / C voxels at a time with consecutive threads on
// consecutive (voxel, channel) elements, into part[b][co]; then the blocks'
// sums are added (a fixed tree per channel).  (C <= 256.)
__global__ void __launch_bounds__(256) k_bias_part(const bf16* __restrict__ dy, long nv, int C, long chunk,
                                                   float* __restrict__ part) {
  __shared__ float red[256];
  const int t = threadIdx.x, G = 256 / C, g = t / C, co = t % C;
  float s = 0.f;
  if (g < G) {
    const long v0 = (long)blockIdx.x * chunk, v1 = min(v0 + chunk, nv);
    for (long v = v0 + g; v < v1; v += G) s += bf2f(dy[v * C + co]);
  }
  red[t] = s;
  __syncthreads();
  if (t < C) {
    float a = 0.f;
    for (int k = 0; k < G; k++) a += red[k * C + t];
    part[(long)blockIdx.x * C + t] = a;
  }
}

// block co: the nb block sums of channel co, 256 at a time, then a tree
__global__ void __launch_bounds__(256) k_bias_sum(const float* __restrict__ part, int nb, int C,
                                                  float* __restrict__ db) {
  __shared__ float red[256];
  const int co = blockIdx.x, t = threadIdx.x;
  float a = 0.f;
  for (int b = t; b < nb; b += 256) a += part[(long)b * C + co];
  red[t] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  if (t == 0) db[co] = red[0];
}

// bf16 weights for the forward (w[co][Kpad]) and input-gradient
// (wt[ci][KpadT]) kernels from FP32 master weights m[co][27*Cin]
__global__ void k_prep_weights(const float* __restrict__ m, int Cin, int Cout, int Kpad, int KpadT,
                               bf16* __restrict__ w, bf16* __restrict__ wt) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nw = (long)Cout * Kpad, nt = (long)Cin * KpadT;
  if (i < nw) {
    const int co = (int)(i / Kpad), k = (int)(i % Kpad);
    w[i] = k < 27 * Cin ? f2bf(m[(long)co * 27 * Cin + k]) : (bf16)0;
  }
  if (i < nt) {
    const int ci = (int)(i / KpadT), k = (int)(i % KpadT);
    bf16 v = 0;
    if (k < 27 * Cout) {
      const int tp = k / Cout, co = k % Cout;
      v = f2bf(m[(long)co * 27 * Cin + (26 - tp) * Cin + ci]);
    }
    wt[i] = v;
  }
}

__global__ void k_pool2(const bf16* __restrict__ in, int Z, int Y, int X, int C, bf16* __restrict__ out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int Z2 = Z / 2, Y2 = Y / 2, X2 = X / 2;
  const long n = (long)Z2 * Y2 * X2 * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long v = i / C;
  const int x = (int)(v % X2), y = (int)((v / X2) % Y2), z = (int)(v / ((long)X2 * Y2));
  float s = 0.f;
  for (int q = 0; q < 8; q++) {
    const int zz = 2 * z + (q >> 2), yy = 2 * y + ((q >> 1) & 1), xx = 2 * x + (q & 1);
    s += bf2f(in[(((long)zz * Y + yy) * X + xx) * C + c]);
  }
  out[i] = f2bf(0.125f * s);
}

__global__ void k_pool2_bwd(const bf16* __restrict__ dout, int Z, int Y, int X, int C, bf16* __restrict__ din,
                            int accumulate) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long n = (long)Z * Y * X * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  long v = i / C;
  const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((long)X * Y));
  const int Y2 = Y / 2, X2 = X / 2;
  float g = 0.125f * bf2f(dout[(((long)(z / 2) * Y2 + y / 2) * X2 + x / 2) * C + c]);
  if (accumulate) g += bf2f(din[i]);
  din[i] = f2bf(g);
}

// nearest 2x upsampling of `in` [Z/2][Y/2][X/2][C1] into channels [0, C1) of
// out [Z][Y][X][C1+C2]; channels [C1, C1+C2) take the skip tensor
__global__ void k_up2_cat(const bf16* __restrict__ in, const bf16* __restrict__ skip, int Z, int Y, int X, int C1,
                          int C2, bf16* __restrict__ out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int C = C1 + C2;
  const long n = (long)Z * Y * X * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const long v = i / C;
  if (c < C1) {
    const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((long)X * Y));
    out[i] = in[(((long)(z / 2) * (Y / 2) + y / 2) * (X / 2) + x / 2) * C1 + c];
  } else {
    out[i] = skip[v * C2 + (c - C1)];
  }
}

// backward of k_up2_cat: din[parent][c] = sum over the 8 children; dskip is
// the gradient of the skip channels (added to dskip_acc when accumulate)
__global__ void k_up2_cat_bwd(const bf16* __restrict__ dout, int Z, int Y, int X, int C1, int C2,
                              bf16* __restrict__ din, bf16* __restrict__ dskip, int accumulate) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int C = C1 + C2;
  const int Z2 = Z / 2, Y2 = Y / 2, X2 = X / 2;
  const long n1 = (long)Z2 * Y2 * X2 * C1, n2 = (long)Z * Y * X * C2;
  if (i < n1) {
    const int c = (int)(i % C1);
    const long v = i / C1;
    const int x = (int)(v % X2), y = (int)((v / X2) % Y2), z = (int)(v / ((long)X2 * Y2));
    float s = 0.f;
    for (int q = 0; q < 8; q++) {
      const int zz = 2 * z + (q >> 2), yy = 2 * y + ((q >> 1) & 1), xx = 2 * x + (q & 1);
      s += bf2f(dout[(((long)zz * Y + yy) * X + xx) * C + c]);
    }
    din[i] = f2bf(s);
  }
  if (i < n2) {
    const int c = (int)(i % C2);
    const long v = i / C2;
    float g = bf2f(dout[v * C + C1 + c]);
    if (accumulate) g += bf2f(dskip[i]);
    dskip[i] = f2bf(g);
  }
}

// g = g * (a > 0 ? 1 : slope): gradient through a LeakyReLU given its output
__global__ void k_dact(bf16* __restrict__ g, const bf16* __restrict__ a, long n, float slope) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float av = bf2f(a[i]);
  g[i] = f2bf(av > 0.f ? bf2f(g[i]) : slope * bf2f(g[i]));
}

// masked mean-square error between out [V][Cs] (FP32) and target [V][Cs];
// grad [V][16] BF16 = 2 (out - t) m / count (channels >= Cs zero); per-block
// partial sums of the squared error
__global__ void k_mse(const float* __restrict__ out, const float* __restrict__ tgt, const float* __restrict__ mask,
                      long nv, int Cs, float inv_count, bf16* __restrict__ grad, float* __restrict__ part) {
  __shared__ float red[256];
  const long v = blockIdx.x * (long)blockDim.x + threadIdx.x;
  float s = 0.f;
  if (v < nv) {
    for (int c = 0; c < 16; c++) {
      float g = 0.f;
      if (c < Cs) {
        const float m = mask[v * Cs + c];
        const float d = (out[v * Cs + c] - tgt[v * Cs + c]) * m;
        s += d * d;
        g = 2.f * d * inv_count;
      }
      grad[v * 16 + c] = f2bf(g);
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long n, float lr, float b1, float b2, float eps, float c1, float c2,
                       float clip) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  gi = fminf(fmaxf(gi, -clip), clip);
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// FP32 [V][C] -> BF16 [V][Cp] (channels >= C zero)
__global__ void k_pack_bf16(const float* __restrict__ in, long nv, int C, int Cp, bf16* __restrict__ out) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= nv * Cp) return;
  const int c = (int)(i % Cp);
  const long v = i / Cp;
  out[i] = c < C ? f2bf(in[v * C + c]) : (bf16)0;
}

// The network refines a field x1 (the field after one multigrid cycle): per
// component a it outputs three channels, a change of log-amplitude dL and a
// relative phasor (1 + q1) + i q2, and the refined field is
//   E_a = x1_a exp(dL) ((1 + q1) + i q2) / |(1 + q1) + i q2|.
// Zero outputs reproduce x1 exactly.

// Network output [n3][n2][n1][9] and x1 -> complex double edge vector x0
// [3][n3+1][n2+1][n1+1] (interior edges; others 0)
__global__ void k_out_to_edges(const float* __restrict__ o, const double2* __restrict__ x1, int n1, int n2, int n3,
                               double2* __restrict__ x0) {
  const long nc = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= 3 * nc) return;
  const int a = (int)(t / nc);
  const long e = t % nc;
  const int i = (int)(e % (n1 + 1)), j = (int)((e / (n1 + 1)) % (n2 + 1)), k = (int)(e / ((long)(n1 + 1) * (n2 + 1)));
  const int p[3] = {i, j, k};
  const int n[3] = {n1, n2, n3};
  const int b = (a + 1) % 3, c = (a + 2) % 3;
  double2 v = make_double2(0.0, 0.0);
  if (p[a] < n[a] && p[b] >= 1 && p[b] <= n[b] - 1 && p[c] >= 1 && p[c] <= n[c] - 1) {
    const long cell = ((long)k * n2 + j) * n1 + i;
    const double dL = fmin(fmax((double)o[cell * 9 + 3 * a], -20.0), 20.0);
    const double qr = 1.0 + (double)o[cell * 9 + 3 * a + 1], qi = o[cell * 9 + 3 * a + 2];
    const double nrm = sqrt(qr * qr + qi * qi);
    const double2 u = x1[t];
    if (nrm > 1e-6) {
      const double s = exp(dL) / nrm;
      v.x = s * (u.x * qr - u.y * qi);
      v.y = s * (u.x * qi + u.y * qr);
    }
  }
  x0[t] = v;
}

// The same with the correction of each edge averaged over the four cells
// that share the edge, so the guess varies as smoothly as x1 between edges;
// cells whose conductivity s11 is below sig_air (the air) take no correction
// (s11 may be null: no cell is excluded).
__global__ void k_out_to_edges_avg(const float* __restrict__ o, const double2* __restrict__ x1,
                                   const float* __restrict__ s11, float sig_air, int n1, int n2, int n3,
                                   double2* __restrict__ x0) {
  const long nc = (long)(n1 + 1) * (n2 + 1) * (n3 + 1);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= 3 * nc) return;
  const int a = (int)(t / nc);
  const long e = t % nc;
  const int i = (int)(e % (n1 + 1)), j = (int)((e / (n1 + 1)) % (n2 + 1)), k = (int)(e / ((long)(n1 + 1) * (n2 + 1)));
  const int p[3] = {i, j, k};
  const int n[3] = {n1, n2, n3};
  const int b = (a + 1) % 3, c = (a + 2) % 3;
  double2 v = make_double2(0.0, 0.0);
  if (p[a] < n[a] && p[b] >= 1 && p[b] <= n[b] - 1 && p[c] >= 1 && p[c] <= n[c] - 1) {
    double dL = 0.0, qr = 0.0, qi = 0.0;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      int q[3] = {p[0], p[1], p[2]};
      q[b] -= s & 1;
      q[c] -= s >> 1;
      const long cell = ((long)q[2] * n2 + q[1]) * n1 + q[0];
      if (s11 && s11[cell] < sig_air) continue;
      dL += 0.25 * fmin(fmax((double)o[cell * 9 + 3 * a], -20.0), 20.0);
      qr += 0.25 * (double)o[cell * 9 + 3 * a + 1];
      qi += 0.25 * (double)o[cell * 9 + 3 * a + 2];
    }
    qr += 1.0;
    const double nrm = sqrt(qr * qr + qi * qi);
    const double2 u = x1[t];
    if (nrm > 1e-6) {
      const double sc = exp(dL) / nrm;
      v.x = sc * (u.x * qr - u.y * qi);
      v.y = sc * (u.x * qi + u.y * qr);
    }
  }
  x0[t] = v;
}

// Training targets [V][9] for refining x1 towards the solved field E, and the
// loss weights: 1 on dL, wp on the phasor channels, where both |E_a| and
// |x1_a| exceed eref (elsewhere 0: the field is negligible there)
__global__ void k_edges_to_target(const double2* __restrict__ E, const double2* __restrict__ x1, int n1, int n2,
                                  int n3, double eref, float wp, float* __restrict__ tgt, float* __restrict__ mask) {
  const long ncell = (long)n1 * n2 * n3;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= ncell) return;
  const int i = (int)(t % n1), j = (int)((t / n1) % n2), k = (int)(t / ((long)n1 * n2));
  const int p[3] = {i, j, k};
  const int n[3] = {n1, n2, n3};
  const long es1 = n1 + 1, es2 = (long)(n1 + 1) * (n2 + 1), nc = es2 * (n3 + 1);
  for (int a = 0; a < 3; a++) {
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const bool in = p[b] >= 1 && p[b] <= n[b] - 1 && p[c] >= 1 && p[c] <= n[c] - 1;
    const long o = a * nc + i + es1 * j + es2 * k;
    const double2 v = in ? E[o] : make_double2(0.0, 0.0);
    const double2 u = in ? x1[o] : make_double2(0.0, 0.0);
    const double av = sqrt(v.x * v.x + v.y * v.y), au = sqrt(u.x * u.x + u.y * u.y);
    const bool ok = in && av > eref && au > eref;
    // relative phasor (E/|E|) conj(x1/|x1|)
    const double rr = ok ? (v.x * u.x + v.y * u.y) / (av * au) : 1.0;
    const double ri = ok ? (v.y * u.x - v.x * u.y) / (av * au) : 0.0;
    tgt[t * 9 + 3 * a] = ok ? (float)log(av / au) : 0.f;
    tgt[t * 9 + 3 * a + 1] = (float)(rr - 1.0);
    tgt[t * 9 + 3 * a + 2] = (float)ri;
    mask[t * 9 + 3 * a] = ok ? 1.f : 0.f;
    mask[t * 9 + 3 * a + 1] = ok ? wp : 0.f;
    mask[t * 9 + 3 * a + 2] = ok ? wp : 0.f;
  }
}

// Input features on the cells [n3][n2][n1][32]:
//   0 log10 sigma11, 1 log10 sigma33, 2..4 (x - xs)/delta per axis (clamped
//   to +-8), 5 log10(r/delta + 1e-3), 6 log10(cell size / delta),
//   7 source blob exp(-(r/h)^2) with h the source cell size, 8 constant 1,
//   9..17 when x1 is given: per component a of the field x1 after one
//   multigrid cycle, ln(max(|x1_a|, eref)/eref)/10 and its unit phasor
__global__ void k_features(const float* __restrict__ s11, const float* __restrict__ s33, const float* __restrict__ xs1,
                           const float* __restrict__ xs2, const float* __restrict__ xs3, const float* __restrict__ d1,
                           const float* __restrict__ d2, const float* __restrict__ d3, int n1, int n2, int n3,
                           float sx, float sy, float sz, float delta, float hsrc, const double2* __restrict__ x1,
                           double eref, bf16* __restrict__ feat) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nc = (long)n1 * n2 * n3;
  if (t >= nc) return;
  const int i = (int)(t % n1), j = (int)((t / n1) % n2), k = (int)(t / ((long)n1 * n2));
  float f[32];
  for (int c = 0; c < 32; c++) f[c] = 0.f;
  f[0] = log10f(fmaxf(s11[t], 1e-12f));
  f[1] = log10f(fmaxf(s33[t], 1e-12f));
  const float dx = xs1[i] - sx, dy = xs2[j] - sy, dz = xs3[k] - sz;
  f[2] = fminf(fmaxf(dx / delta, -8.f), 8.f);
  f[3] = fminf(fmaxf(dy / delta, -8.f), 8.f);
  f[4] = fminf(fmaxf(dz / delta, -8.f), 8.f);
  const float r = sqrtf(dx * dx + dy * dy + dz * dz);
  f[5] = log10f(r / delta + 1e-3f);
  f[6] = log10f(cbrtf(d1[i] * d2[j] * d3[k]) / delta);
  f[7] = expf(-(r / hsrc) * (r / hsrc));
  f[8] = 1.f;
  if (x1) {
    const int p[3] = {i, j, k};
    const int n[3] = {n1, n2, n3};
    const long es1 = n1 + 1, es2 = (long)(n1 + 1) * (n2 + 1), ncc = es2 * (n3 + 1);
    for (int a = 0; a < 3; a++) {
      const int b = (a + 1) % 3, c = (a + 2) % 3;
      const bool in = p[b] >= 1 && p[b] <= n[b] - 1 && p[c] >= 1 && p[c] <= n[c] - 1;
      const double2 v = in ? x1[a * ncc + i + es1 * j + es2 * k] : make_double2(0.0, 0.0);
      const double amp = sqrt(v.x * v.x + v.y * v.y);
      f[9 + 3 * a] = (float)(0.1 * log(fmax(amp, eref) / eref));
      f[10 + 3 * a] = amp > eref ? (float)(v.x / amp) : 0.f;
      f[11 + 3 * a] = amp > eref ? (float)(v.y / amp) : 0.f;
    }
  }
  for (int c = 0; c < 32; c++) feat[t * 32 + c] = f2bf(f[c]);
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, long n, uint32_t v) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_scale(double2* __restrict__ v, long n, double s) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) { v[i].x *= s; v[i].y *= s; }
}

inline unsigned blocks(long n, int b) { return (unsigned)((n + b - 1) / b); }
hipStream_t S(void* s) { return (hipStream_t)s; }

template <int MT>
void conv_fwd_mt(const bf16* x, int Z, int Y, int X, int Cin, const bf16* w, int Kpad, int Cout, const float* bias,
                 int epi, float slope, const bf16* act, bf16* ob, float* of, int cstore, hipStream_t s) {
  // the last tile of a row may be partial: loads and stores check x < X
  const dim3 g((X + 16 * MT - 1) / (16 * MT), (Y + 3) / 4, Z), b(256);
  hipEvent_t e0, e1;
  emmi::timing::kernel_events(emmi::timing::CONV, 54.0 * Cin * Cout * (double)Z * Y * X, &e0, &e1);
  switch (Cout / 16) {
#define CASE(NT)                                                                                                   \
  case NT:                                                                                                         \
    if (e0) hipExtLaunchKernelGGL((k_conv_fwd<MT, NT>), g, b, 0, s, e0, e1, 0, x, Z, Y, X, Cin, w, Kpad, bias, epi, \
                                  slope, act, ob, of, cstore);                                                     \
    else hipLaunchKernelGGL((k_conv_fwd<MT, NT>), g, b, 0, s, x, Z, Y, X, Cin, w, Kpad, bias, epi, slope, act, ob,  \
                            of, cstore);                                                                           \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
  }
}

}  // namespace

extern "C" {

int emmi_x0_kpad(int cin) { return (27 * cin + 31) / 32 * 32; }

int emmi_x0_conv_fwd(const void* x, int Z, int Y, int X, int Cin, const void* w, int Cout, const float* bias, int epi,
                     float slope, const void* act, void* out_bf16, float* out_f32, int cout_store, void* stream) {
  if (Cin % 16 || Cout % 16 || Cout > 128 || Cin <= 0 || X <= 0 || Y <= 0 || Z <= 0) return -1;
  const int Kpad = emmi_x0_kpad(Cin);
  const bf16* xb = (const bf16*)x;
  const bf16* wb = (const bf16*)w;
  if (X % 64 == 0) conv_fwd_mt<4>(xb, Z, Y, X, Cin, wb, Kpad, Cout, bias, epi, slope, (const bf16*)act, (bf16*)out_bf16, out_f32, cout_store, S(stream));
  else if (X % 32 == 0) conv_fwd_mt<2>(xb, Z, Y, X, Cin, wb, Kpad, Cout, bias, epi, slope, (const bf16*)act, (bf16*)out_bf16, out_f32, cout_store, S(stream));
  else conv_fwd_mt<1>(xb, Z, Y, X, Cin, wb, Kpad, Cout, bias, epi, slope, (const bf16*)act, (bf16*)out_bf16, out_f32, cout_store, S(stream));
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_conv_dw(const void* x, const void* dy, int Z, int Y, int X, int Cin, int Cout, float* part, int zslab,
                    float* dw, float* db, void* stream) {
  if (Cin % 16 || Cout % 16 || Cout > 256 || zslab < 1 || Z < 1 || Y < 1 || X < 1) return -1;
  if ((long)Z * Y * X >= (1L << 31)) return -1;   // the kernel indexes voxels in 32 bits
  const int nslab = (Z + zslab - 1) / zslab;
  const dim3 g(27 * (Cin / 16) * (Cout / 16), nslab), b(256);
  hipEvent_t e0, e1;
  emmi::timing::kernel_events(emmi::timing::CONV_DW, 54.0 * Cin * Cout * (double)Z * Y * X, &e0, &e1);
  if (e0) hipExtLaunchKernelGGL(k_conv_dw, g, b, 0, S(stream), e0, e1, 0, (const bf16*)x, (const bf16*)dy, Z, Y, X, Cin,
                                Cout, zslab, part);
  else hipLaunchKernelGGL(k_conv_dw, g, b, 0, S(stream), (const bf16*)x, (const bf16*)dy, Z, Y, X, Cin, Cout, zslab, part);
  const long n = 27L * Cin * Cout;
  hipLaunchKernelGGL(k_reduce_slabs, dim3(blocks(n, 256)), dim3(256), 0, S(stream), part, nslab, n, dw);
  if (db) {
    // the bias partials reuse part once the slab reduction has read it (same
    // stream): nb Cout floats, nb <= nslab 27 Cin
    const long nv = (long)Z * Y * X;
    const long nbl = std::min<long>(std::min<long>(1024, (long)nslab * 27 * Cin), nv);
    const long chunk = (nv + nbl - 1) / nbl;
    const int nb = (int)((nv + chunk - 1) / chunk);
    hipLaunchKernelGGL(k_bias_part, dim3(nb), dim3(256), 0, S(stream), (const bf16*)dy, nv, Cout, chunk, part);
    hipLaunchKernelGGL(k_bias_sum, dim3(Cout), dim3(256), 0, S(stream), (const float*)part, nb, Cout, db);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_prep_weights(const float* master, int Cin, int Cout, void* w, void* wt, void* stream) {
  const int Kpad = emmi_x0_kpad(Cin), KpadT = emmi_x0_kpad(Cout);
  const long n = std::max((long)Cout * Kpad, (long)Cin * KpadT);
  hipLaunchKernelGGL(k_prep_weights, dim3(blocks(n, 256)), dim3(256), 0, S(stream), master, Cin, Cout, Kpad, KpadT,
                     (bf16*)w, (bf16*)wt);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_pool2(const void* in, int Z, int Y, int X, int C, void* out, void* stream) {
  const long n = (long)(Z / 2) * (Y / 2) * (X / 2) * C;
  hipLaunchKernelGGL(k_pool2, dim3(blocks(n, 256)), dim3(256), 0, S(stream), (const bf16*)in, Z, Y, X, C, (bf16*)out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_pool2_bwd(const void* dout, int Z, int Y, int X, int C, void* din, int accumulate, void* stream) {
  const long n = (long)Z * Y * X * C;
  hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks(n, 256)), dim3(256), 0, S(stream), (const bf16*)dout, Z, Y, X, C, (bf16*)din,
                     accumulate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_up2_cat(const void* in, const void* skip, int Z, int Y, int X, int C1, int C2, void* out, void* stream) {
  const long n = (long)Z * Y * X * (C1 + C2);
  hipLaunchKernelGGL(k_up2_cat, dim3(blocks(n, 256)), dim3(256), 0, S(stream), (const bf16*)in, (const bf16*)skip, Z, Y,
                     X, C1, C2, (bf16*)out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_up2_cat_bwd(const void* dout, int Z, int Y, int X, int C1, int C2, void* din, void* dskip, int accumulate,
                        void* stream) {
  const long n = std::max((long)(Z / 2) * (Y / 2) * (X / 2) * C1, (long)Z * Y * X * C2);
  hipLaunchKernelGGL(k_up2_cat_bwd, dim3(blocks(n, 256)), dim3(256), 0, S(stream), (const bf16*)dout, Z, Y, X, C1, C2,
                     (bf16*)din, (bf16*)dskip, accumulate);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_dact(void* g, const void* a, long n, float slope, void* stream) {
  hipLaunchKernelGGL(k_dact, dim3(blocks(n, 256)), dim3(256), 0, S(stream), (bf16*)g, (const bf16*)a, n, slope);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_mse(const float* out, const float* tgt, const float* mask, long nv, int cs, float inv_count, void* grad,
                float* part, void* stream) {
  hipLaunchKernelGGL(k_mse, dim3(blocks(nv, 256)), dim3(256), 0, S(stream), out, tgt, mask, nv, cs, inv_count,
                     (bf16*)grad, part);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_adam(float* p, const float* g, float* m1, float* m2, long n, float lr, float b1, float b2, float eps,
                 int step, float clip, void* stream) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, S(stream), p, g, m1, m2, n, lr, b1, b2, eps, c1, c2,
                     clip);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_pack_bf16(const float* in, long nv, int C, int Cp, void* out, void* stream) {
  hipLaunchKernelGGL(k_pack_bf16, dim3(blocks(nv * Cp, 256)), dim3(256), 0, S(stream), in, nv, C, Cp, (bf16*)out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_out_to_edges(const float* o, const void* x1, int n1, int n2, int n3, void* x0, void* stream) {
  const long n = 3L * (n1 + 1) * (n2 + 1) * (n3 + 1);
  hipLaunchKernelGGL(k_out_to_edges, dim3(blocks(n, 256)), dim3(256), 0, S(stream), o, (const double2*)x1, n1, n2, n3,
                     (double2*)x0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_out_to_edges_avg(const float* o, const void* x1, const float* s11, float sig_air, int n1, int n2, int n3,
                             void* x0, void* stream) {
  const long n = 3L * (n1 + 1) * (n2 + 1) * (n3 + 1);
  hipLaunchKernelGGL(k_out_to_edges_avg, dim3(blocks(n, 256)), dim3(256), 0, S(stream), o, (const double2*)x1, s11,
                     sig_air, n1, n2, n3, (double2*)x0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_edges_to_target(const void* E, const void* x1, int n1, int n2, int n3, double eref, float wp, float* tgt,
                            float* mask, void* stream) {
  const long n = (long)n1 * n2 * n3;
  hipLaunchKernelGGL(k_edges_to_target, dim3(blocks(n, 256)), dim3(256), 0, S(stream), (const double2*)E,
                     (const double2*)x1, n1, n2, n3, eref, wp, tgt, mask);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_features(const float* s11, const float* s33, const float* xs1, const float* xs2, const float* xs3,
                     const float* d1, const float* d2, const float* d3, int n1, int n2, int n3, float sx, float sy,
                     float sz, float delta, float hsrc, const void* x1, double eref, void* feat, void* stream) {
  const long n = (long)n1 * n2 * n3;
  hipLaunchKernelGGL(k_features, dim3(blocks(n, 256)), dim3(256), 0, S(stream), s11, s33, xs1, xs2, xs3, d1, d2, d3,
                     n1, n2, n3, sx, sy, sz, delta, hsrc, (const double2*)x1, eref, (bf16*)feat);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_fill(void* p, long nwords, unsigned value, void* stream) {
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks(nwords, 256)), dim3(256), 0, S(stream), (uint32_t*)p, nwords, value);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int emmi_x0_scale(void* v, long n, double s, void* stream) {
  hipLaunchKernelGGL(k_scale, dim3(blocks(n, 256)), dim3(256), 0, S(stream), (double2*)v, n, s);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"
