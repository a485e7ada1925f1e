// Learned coarse correction (no reference counterpart): a message-passing
// network over one coarse AMG level's graph that maps a restricted residual
// to a coarse correction, run and trained on the device.
//
// Graph: the level operator A (symmetric, positive diagonal). Node i's
// neighbours are the off-diagonal entries of row i, weighted
// S_ij = |a_ij| / a_ii in the forward pass; the backward pass gathers through
// S^T, whose row i holds S_ji = |a_ij| / a_jj (A is symmetric), so no kernel
// scatters.
//
// Network (hidden width 64, L message-passing layers, 4 input features):
//   x_i = (r_i / rms(r), a_ii / mean diag, sum_j |a_ij| / mean diag, 1)
//   h0  = relu(x W0 + b0)
//   m_l = S h_{l-1};  h_l = relu(h_{l-1} Ws_l + m_l Wn_l + b_l)
//   y   = h_L wo + bo;  correction e = y * rms(r) / mean diag
// Training minimises the mean squared error of y against the exact
// correction in the same units, with Adam.
//
// Precision: the layer contractions run on the bf16 MFMA
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation; activations h_l and the
// aggregates m_l are stored bf16, master weights, Adam moments, the input and
// output layers, backpropagated gradients of h and m, and the bias gradients
// stay FP32, and every sum over nodes is deterministic: each kernel that
// contributes to a parameter gradient runs kLearnedSlots workgroups over
// strided tiles and stores one FP32 partial per slot; the update kernel sums
// the slots in FP64 in a fixed order.
//
// Parameter vector (FP32, row major): W0 (4 x 64), b0 (64); per layer
// Ws (64 x 64), Wn (64 x 64), b (64); wo (64), bo. Packed bf16 copies per
// layer: wf = [Ws; Wn]^T as 64 rows of 128 (the forward B operand, one
// output column per row) and wb = [Ws; Wn] as 128 rows of 64 (the backward
// B operand).
#include "kh_device_math.hpp"

namespace kh {
namespace {

using bf16_t = unsigned short;
typedef short bf16x8 __attribute__((__vector_size__(8 * sizeof(short))));
typedef float f32x16 __attribute__((__vector_size__(16 * sizeof(float))));

constexpr int kH = KH_LEARNED_HIDDEN;  // hidden width; the tiling assumes 64
constexpr int kF = 4;                  // input features
constexpr int kTile = 64;              // nodes per tile
constexpr int kSlots = KH_LEARNED_SLOTS;
constexpr int kXs = 2 * kH + 8;        // LDS row stride (bf16) of a 128-wide row, padded
constexpr int kZs = kH + 8;            // LDS row stride (bf16) of a 64-wide row, padded
constexpr int kLayerParams = 2 * kH * kH + kH;
constexpr int kInputParams = kF * kH + kH;

static_assert(kH == 64, "the MFMA tiling assumes a hidden width of 64");

// Round to nearest even (NaN stays NaN).
__device__ __forceinline__ bf16_t to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return bf16_t((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return bf16_t(u >> 16);
}
__device__ __forceinline__ float from_bf16(bf16_t h) { return __uint_as_float(unsigned(h) << 16); }
__device__ __forceinline__ unsigned pack2(float a, float b) {
  return unsigned(to_bf16(a)) | (unsigned(to_bf16(b)) << 16);
}
__device__ __forceinline__ float lo_bf16(unsigned w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float hi_bf16(unsigned w) { return __uint_as_float(w & 0xffff0000u); }

// Eight bf16 values of a 16-byte word, as floats.
__device__ __forceinline__ void unpack8(uint4 u, float* out) {
  out[0] = lo_bf16(u.x); out[1] = hi_bf16(u.x);
  out[2] = lo_bf16(u.y); out[3] = hi_bf16(u.y);
  out[4] = lo_bf16(u.z); out[5] = hi_bf16(u.z);
  out[6] = lo_bf16(u.w); out[7] = hi_bf16(u.w);
}
__device__ __forceinline__ uint4 pack8(const float* v) {
  return make_uint4(pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7]));
}

// acc += A[r0:r0+32, 0:K] * B[0:K, c0:c0+32] on one wave. A is row major in
// LDS (stride sa); B is held transposed (row c of bt is column c of B, stride
// sb). Lane l supplies A[r0 + l%32][k + 8(l/32) + v] and B[k + 8(l/32) + v]
// [c0 + l%32], v = 0..7, to v_mfma_f32_32x32x16_bf16.
__device__ __forceinline__ void mfma_block(f32x16& acc, const bf16_t* la, int sa, const bf16_t* bt,
                                           int sb, int r0, int c0, int K) {
  const int lane = threadIdx.x & 63;
  const bf16_t* pa = la + (r0 + (lane & 31)) * sa + 8 * (lane >> 5);
  const bf16_t* pb = bt + (c0 + (lane & 31)) * sb + 8 * (lane >> 5);
  for (int k = 0; k < K; k += 16) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(pa + k);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(pb + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
}
// Accumulator register q of a lane holds row d_row(q) and column lane % 32 of
// the 32 x 32 block.
__device__ __forceinline__ int d_row(int q) {
  return 8 * (q >> 2) + 4 * ((threadIdx.x & 63) >> 5) + (q & 3);
}

// Sum of v over the lanes of a wave whose lane index agrees in bits 0..1
// (the 16 lanes that hold the same column group), valid in every lane.
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int off = 4; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
  return v;
}
// Reduces count values per thread (thread t owns columns 16 (t%4) + c) over
// the workgroup into out[16 * part + c] for the 64 columns; lds holds
// 4 waves x 64 floats. Valid for threads 0..63 after the call.
template <int Count>
__device__ __forceinline__ void column_reduce(float (&v)[Count], float* lds, float* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, part = threadIdx.x & 3;
  static_assert(Count == 16, "one 16-column group per thread");
#pragma unroll
  for (int c = 0; c < Count; ++c) v[c] = group_sum(v[c]);
  if (lane < 4) {
#pragma unroll
    for (int c = 0; c < Count; ++c) lds[wave * 64 + 16 * part + c] = v[c];
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    float s = 0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += lds[w * 64 + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// ------------------------------------------------------------------ graph
// Raw diagonal and off-diagonal l1 sum per node, and FP64 slot sums of the
// diagonal.
__global__ void __launch_bounds__(kBlock) k_learned_graph_diag(int n, const int* __restrict__ rp,
    const int* __restrict__ col, const float* __restrict__ val, float* __restrict__ nf,
    double* __restrict__ partials) {
  double mine = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    float d = 0, off = 0;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
      if (col[j] == i) d = val[j];
      else off += fabsf(val[j]);
    }
    nf[2 * i] = d;
    nf[2 * i + 1] = off;
    mine += double(d);
  }
  __shared__ double lds[kBlock / 64];
  mine = block_sum(mine, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = mine;
}
// Per-entry neighbour weights: ws = |a_ij| / a_ii (forward), wt = |a_ij| / a_jj
// (transpose); zero on the diagonal.
__global__ void __launch_bounds__(kBlock) k_learned_graph_weights(int n, const int* __restrict__ rp,
    const int* __restrict__ col, const float* __restrict__ val, const float* __restrict__ nf,
    float* __restrict__ ws, float* __restrict__ wt) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float di = nf[2 * i];
  for (int j = rp[i]; j < rp[i + 1]; ++j) {
    const int c = col[j];
    const float a = fabsf(val[j]);
    ws[j] = c == i ? 0.0f : a / di;
    wt[j] = c == i ? 0.0f : a / nf[2 * c];
  }
}
// Node features divided by the mean diagonal; gstats[0] = mean diagonal.
__global__ void __launch_bounds__(kBlock) k_learned_graph_scale(int n, float* __restrict__ nf,
    const double* __restrict__ partials, float* __restrict__ gstats) {
  __shared__ double lds[kBlock / 64];
  __shared__ double total;
  double v = threadIdx.x < kSlots ? partials[threadIdx.x] : 0.0;
  v = block_sum(v, lds);
  if (threadIdx.x == 0) total = v;
  __syncthreads();
  const float mean = float(total / double(n));
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0) gstats[0] = mean;
  if (i >= n) return;
  nf[2 * i] /= mean;
  nf[2 * i + 1] /= mean;
}

// --------------------------------------------------------------- forward
// FP64 slot sums of r^2.
__global__ void __launch_bounds__(kBlock) k_learned_norm(int n, const float* __restrict__ r,
                                                         double* __restrict__ partials) {
  double mine = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
    mine += double(r[i]) * double(r[i]);
  __shared__ double lds[kBlock / 64];
  mine = block_sum(mine, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = mine;
}

// Features and the input layer, one thread per (node, 8 columns); stats[0] =
// rms(r) / mean diagonal (the output scale), stats[1] its inverse (0 when r = 0).
__global__ void __launch_bounds__(kBlock) k_learned_input(int n, const float* __restrict__ r,
    const double* __restrict__ partials, const float* __restrict__ nf,
    const float* __restrict__ gstats, const float* __restrict__ theta, float* __restrict__ x,
    bf16_t* __restrict__ h0, float* __restrict__ stats) {
  __shared__ double lds[kBlock / 64];
  __shared__ double total;
  double s = threadIdx.x < kSlots ? partials[threadIdx.x] : 0.0;
  s = block_sum(s, lds);
  if (threadIdx.x == 0) total = s;
  __syncthreads();
  const double rms = sqrt(total / double(n));
  const float inv_rms = rms > 0 ? float(1.0 / rms) : 0.0f;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double scale = rms / double(gstats[0]);
    stats[0] = float(scale);
    stats[1] = scale > 0 ? float(1.0 / scale) : 0.0f;
  }
  const long idx = long(blockIdx.x) * kBlock + threadIdx.x;
  const int i = int(idx >> 3), c8 = int(idx & 7) * 8;
  if (i >= n) return;
  const float f[kF] = {r[i] * inv_rms, nf[2 * i], nf[2 * i + 1], 1.0f};
  if (c8 == 0) *reinterpret_cast<float4*>(&x[4 * long(i)]) = make_float4(f[0], f[1], f[2], f[3]);
  float h[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    float z = theta[kF * kH + c8 + c];
#pragma unroll
    for (int k = 0; k < kF; ++k) z += f[k] * theta[k * kH + c8 + c];
    h[c] = fmaxf(z, 0.0f);
  }
  *reinterpret_cast<uint4*>(&h0[long(i) * kH + c8]) = pack8(h);
}

// One message-passing layer over a 64-node tile: the tile's rows of h_{l-1}
// and their aggregates m = S h_{l-1} (gathered in FP32, stored bf16) form the
// 64 x 128 A operand; [Ws; Wn] the 128 x 64 B operand; each wave computes a
// 32 x 32 block of h_l = relu(A B + b).
__global__ void __launch_bounds__(kBlock) k_learned_layer(int n, const int* __restrict__ rp,
    const int* __restrict__ col, const float* __restrict__ ws, const bf16_t* __restrict__ hprev,
    const bf16_t* __restrict__ wf, const float* __restrict__ bias, bf16_t* __restrict__ hout,
    bf16_t* __restrict__ mout) {
  __shared__ __align__(16) bf16_t X[kTile * kXs];
  __shared__ __align__(16) bf16_t W[kH * kXs];
  const int t = threadIdx.x;
  const long node0 = long(blockIdx.x) * kTile;
  for (int q = t; q < kH * 16; q += kBlock) {
    const int c = q >> 4, k8 = (q & 15) * 8;
    *reinterpret_cast<uint4*>(&W[c * kXs + k8]) = *reinterpret_cast<const uint4*>(&wf[c * 2 * kH + k8]);
  }
  for (int q = t; q < kTile * 8; q += kBlock) {
    const int row = q >> 3, k8 = (q & 7) * 8;
    const long i = node0 + row;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (i < n) v = *reinterpret_cast<const uint4*>(&hprev[i * kH + k8]);
    *reinterpret_cast<uint4*>(&X[row * kXs + k8]) = v;
  }
  {
    const int row = t >> 2, c0 = (t & 3) * 16;
    const long i = node0 + row;
    float acc[16] = {};
    if (i < n) {
      for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const float w = ws[j];
        if (w == 0.0f) continue;
        const uint4* src = reinterpret_cast<const uint4*>(&hprev[long(col[j]) * kH + c0]);
        float hv[16];
        unpack8(src[0], hv);
        unpack8(src[1], hv + 8);
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[c] += w * hv[c];
      }
    }
    const uint4 p0 = pack8(acc), p1 = pack8(acc + 8);
    *reinterpret_cast<uint4*>(&X[row * kXs + kH + c0]) = p0;
    *reinterpret_cast<uint4*>(&X[row * kXs + kH + c0 + 8]) = p1;
    if (i < n) {
      *reinterpret_cast<uint4*>(&mout[i * kH + c0]) = p0;
      *reinterpret_cast<uint4*>(&mout[i * kH + c0 + 8]) = p1;
    }
  }
  __syncthreads();
  const int wave = t >> 6, lane = t & 63;
  const int r0 = 32 * (wave >> 1), c0 = 32 * (wave & 1);
  f32x16 acc = {};
  mfma_block(acc, X, kXs, W, kXs, r0, c0, 2 * kH);
  const int c = c0 + (lane & 31);
  const float bc = bias[c];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const long i = node0 + r0 + d_row(q);
    if (i < n) hout[i * kH + c] = to_bf16(fmaxf(acc[q] + bc, 0.0f));
  }
}

// Output layer for inference: e_i = (h_L wo + bo) * scale.
__global__ void __launch_bounds__(kBlock) k_learned_head(int n, const bf16_t* __restrict__ hl,
    const float* __restrict__ head, const float* __restrict__ stats, float* __restrict__ e) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint4* src = reinterpret_cast<const uint4*>(&hl[long(i) * kH]);
  float y = head[kH];
  for (int q = 0; q < kH / 8; ++q) {
    float hv[8];
    unpack8(src[q], hv);
#pragma unroll
    for (int c = 0; c < 8; ++c) y += hv[c] * head[8 * q + c];
  }
  e[i] = y * stats[0];
}

// Output layer and loss for training, kLearnedSlots workgroups over strided
// nodes: y, dy = 2 (y - target / scale) / n, and per slot the partial
// gradients of wo and bo, the partial sum of squared errors and the partial
// sum of squared normalised targets at partials[slot * width + 0 .. 66]
// (wo, bo, loss, target energy).
__global__ void __launch_bounds__(kBlock) k_learned_head_loss(int n, const bf16_t* __restrict__ hl,
    const float* __restrict__ head, const float* __restrict__ stats,
    const float* __restrict__ target, float* __restrict__ y_out, float* __restrict__ dy_out,
    float* __restrict__ partials, int width) {
  const float inv_scale = stats[1];
  float gwo[kH] = {};
  float gbo = 0;
  double loss = 0, energy = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const uint4* src = reinterpret_cast<const uint4*>(&hl[long(i) * kH]);
    float hv[kH];
    for (int q = 0; q < kH / 8; ++q) unpack8(src[q], hv + 8 * q);
    float y = head[kH];
#pragma unroll
    for (int c = 0; c < kH; ++c) y += hv[c] * head[c];
    const float t = target[i] * inv_scale;
    const float diff = y - t;
    const float dy = 2.0f * diff / float(n);
    y_out[i] = y;
    dy_out[i] = dy;
    loss += double(diff) * double(diff);
    energy += double(t) * double(t);
    gbo += dy;
#pragma unroll
    for (int c = 0; c < kH; ++c) gwo[c] += hv[c] * dy;
  }
  __shared__ float lds[kBlock / 64][kH + 2];
  __shared__ double dl[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kH; ++c) {
    float v = gwo[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) lds[wave][c] = v;
  }
  float vb = gbo;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) vb += __shfl_xor(vb, off, 64);
  if (lane == 0) lds[wave][kH] = vb;
  loss = block_sum(loss, dl);  // synchronises the lds writes above
  energy = block_sum(energy, dl);
  if (threadIdx.x <= kH) {
    float s = 0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += lds[w][threadIdx.x];
    partials[long(blockIdx.x) * width + threadIdx.x] = s;
  }
  if (threadIdx.x == 0) {
    partials[long(blockIdx.x) * width + kH + 1] = float(loss);
    partials[long(blockIdx.x) * width + kH + 2] = float(energy);
  }
}

// -------------------------------------------------------------- backward
// Backward through layer l over strided 64-node tiles:
//   dh_l = dy wo^T (last layer) or p + S^T dm (from layer l + 1)
//   dz   = dh_l * (h_l > 0)                        -> dz (bf16), bias partials
//   [p | dm] = dz [Ws^T | Wn^T]                    -> p_out, dm_out (FP32)
// so that dh_{l-1} = p_out + S^T dm_out.
__global__ void __launch_bounds__(kBlock) k_learned_back(int n, int last,
    const float* __restrict__ dy, const float* __restrict__ head, const float* __restrict__ pin,
    const float* __restrict__ dmin, const int* __restrict__ rp, const int* __restrict__ col,
    const float* __restrict__ wt, const bf16_t* __restrict__ hl, const bf16_t* __restrict__ wb,
    bf16_t* __restrict__ dz, float* __restrict__ pout, float* __restrict__ dmout,
    float* __restrict__ partials, int width) {
  __shared__ __align__(16) bf16_t Z[kTile * kZs];
  __shared__ __align__(16) bf16_t B[2 * kH * kZs];
  __shared__ float red[kBlock / 64 * 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int q = t; q < 2 * kH * 8; q += kBlock) {
    const int row = q >> 3, k8 = (q & 7) * 8;
    *reinterpret_cast<uint4*>(&B[row * kZs + k8]) = *reinterpret_cast<const uint4*>(&wb[row * kH + k8]);
  }
  const int row = t >> 2, c0 = (t & 3) * 16;
  float bsum[16] = {};
  const int tiles = (n + kTile - 1) / kTile;
  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const long node0 = long(tile) * kTile, i = node0 + row;
    float d[16] = {};
    if (i < n) {
      if (last) {
        const float g = dy[i];
#pragma unroll
        for (int c = 0; c < 16; ++c) d[c] = g * head[c0 + c];
      } else {
        const float4* p = reinterpret_cast<const float4*>(&pin[i * kH + c0]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float4 v = p[q];
          d[4 * q] = v.x; d[4 * q + 1] = v.y; d[4 * q + 2] = v.z; d[4 * q + 3] = v.w;
        }
        for (int j = rp[i]; j < rp[i + 1]; ++j) {
          const float w = wt[j];
          if (w == 0.0f) continue;
          const float4* m = reinterpret_cast<const float4*>(&dmin[long(col[j]) * kH + c0]);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float4 v = m[q];
            d[4 * q] += w * v.x; d[4 * q + 1] += w * v.y; d[4 * q + 2] += w * v.z; d[4 * q + 3] += w * v.w;
          }
        }
      }
      const uint4* hs = reinterpret_cast<const uint4*>(&hl[i * kH + c0]);
      float hv[16];
      unpack8(hs[0], hv);
      unpack8(hs[1], hv + 8);
#pragma unroll
      for (int c = 0; c < 16; ++c) {
        if (!(hv[c] > 0.0f)) d[c] = 0.0f;
        bsum[c] += d[c];
      }
    }
    const uint4 z0 = pack8(d), z1 = pack8(d + 8);
    *reinterpret_cast<uint4*>(&Z[row * kZs + c0]) = z0;
    *reinterpret_cast<uint4*>(&Z[row * kZs + c0 + 8]) = z1;
    if (i < n) {
      *reinterpret_cast<uint4*>(&dz[i * kH + c0]) = z0;
      *reinterpret_cast<uint4*>(&dz[i * kH + c0 + 8]) = z1;
    }
    __syncthreads();
    const int r0 = 32 * (wave & 1);
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      const int cb = 32 * (2 * (wave >> 1) + half);  // 0, 32, 64, 96
      f32x16 acc = {};
      mfma_block(acc, Z, kZs, B, kZs, r0, cb, kH);
      const int c = cb + (lane & 31);
      float* out = c < kH ? pout : dmout;
      const int cc = c < kH ? c : c - kH;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const long k = node0 + r0 + d_row(q);
        if (k < n) out[k * kH + cc] = acc[q];
      }
    }
    __syncthreads();
  }
  column_reduce(bsum, red, partials + long(blockIdx.x) * width);
}

// Weight gradient of one layer, [Ws; Wn] += [h_{l-1} | m_l]^T dz_l, over
// strided 64-node tiles staged transposed in LDS; wave w accumulates rows
// 32w..32w+31 of the 128 x 64 result, stored per slot at partials[slot * width].
__global__ void __launch_bounds__(kBlock) k_learned_weight_grad(int n, const bf16_t* __restrict__ hprev,
    const bf16_t* __restrict__ m, const bf16_t* __restrict__ dz, float* __restrict__ partials,
    int width) {
  __shared__ __align__(16) bf16_t AT[2 * kH * kZs];
  __shared__ __align__(16) bf16_t DT[kH * kZs];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  f32x16 acc0 = {}, acc1 = {};
  const int tiles = (n + kTile - 1) / kTile;
  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const long node0 = long(tile) * kTile;
    for (int q = t; q < kTile * 16; q += kBlock) {
      const int node = q >> 4, f0 = (q & 15) * 8;
      const long i = node0 + node;
      uint4 u = make_uint4(0, 0, 0, 0);
      if (i < n) u = *reinterpret_cast<const uint4*>(f0 < kH ? &hprev[i * kH + f0] : &m[i * kH + f0 - kH]);
      const bf16_t v[8] = {bf16_t(u.x), bf16_t(u.x >> 16), bf16_t(u.y), bf16_t(u.y >> 16),
                           bf16_t(u.z), bf16_t(u.z >> 16), bf16_t(u.w), bf16_t(u.w >> 16)};
#pragma unroll
      for (int k = 0; k < 8; ++k) AT[(f0 + k) * kZs + node] = v[k];
    }
    for (int q = t; q < kTile * 8; q += kBlock) {
      const int node = q >> 3, c8 = (q & 7) * 8;
      const long i = node0 + node;
      uint4 u = make_uint4(0, 0, 0, 0);
      if (i < n) u = *reinterpret_cast<const uint4*>(&dz[i * kH + c8]);
      const bf16_t v[8] = {bf16_t(u.x), bf16_t(u.x >> 16), bf16_t(u.y), bf16_t(u.y >> 16),
                           bf16_t(u.z), bf16_t(u.z >> 16), bf16_t(u.w), bf16_t(u.w >> 16)};
#pragma unroll
      for (int k = 0; k < 8; ++k) DT[(c8 + k) * kZs + node] = v[k];
    }
    __syncthreads();
    mfma_block(acc0, AT, kZs, DT, kZs, 32 * wave, 0, kTile);
    mfma_block(acc1, AT, kZs, DT, kZs, 32 * wave, 32, kTile);
    __syncthreads();
  }
  float* out = partials + long(blockIdx.x) * width;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int f = 32 * wave + d_row(q);
    out[f * kH + (lane & 31)] = acc0[q];
    out[f * kH + 32 + (lane & 31)] = acc1[q];
  }
}

// Backward through the input layer: dh_0 = p + S^T dm, dz_0 = dh_0 (h_0 > 0),
// per slot the partial gradients of W0 (x^T dz_0) and b0 at
// partials[slot * width + 0 .. 319].
__global__ void __launch_bounds__(kBlock) k_learned_back_input(int n, const float* __restrict__ pin,
    const float* __restrict__ dmin, const int* __restrict__ rp, const int* __restrict__ col,
    const float* __restrict__ wt, const bf16_t* __restrict__ h0, const float* __restrict__ x,
    float* __restrict__ partials, int width) {
  __shared__ float red[kBlock / 64 * 64];
  const int t = threadIdx.x;
  const int row = t >> 2, c0 = (t & 3) * 16;
  float gw[kF][16] = {};
  float gb[16] = {};
  const int tiles = (n + kTile - 1) / kTile;
  for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const long i = long(tile) * kTile + row;
    if (i >= n) continue;
    float d[16];
    const float4* p = reinterpret_cast<const float4*>(&pin[i * kH + c0]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 v = p[q];
      d[4 * q] = v.x; d[4 * q + 1] = v.y; d[4 * q + 2] = v.z; d[4 * q + 3] = v.w;
    }
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
      const float w = wt[j];
      if (w == 0.0f) continue;
      const float4* m = reinterpret_cast<const float4*>(&dmin[long(col[j]) * kH + c0]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 v = m[q];
        d[4 * q] += w * v.x; d[4 * q + 1] += w * v.y; d[4 * q + 2] += w * v.z; d[4 * q + 3] += w * v.w;
      }
    }
    const uint4* hs = reinterpret_cast<const uint4*>(&h0[i * kH + c0]);
    float hv[16];
    unpack8(hs[0], hv);
    unpack8(hs[1], hv + 8);
    const float4 xf = *reinterpret_cast<const float4*>(&x[4 * i]);
    const float xs[kF] = {xf.x, xf.y, xf.z, xf.w};
#pragma unroll
    for (int c = 0; c < 16; ++c) {
      const float z = hv[c] > 0.0f ? d[c] : 0.0f;
      gb[c] += z;
#pragma unroll
      for (int k = 0; k < kF; ++k) gw[k][c] += xs[k] * z;
    }
  }
  float* out = partials + long(blockIdx.x) * width;
#pragma unroll
  for (int k = 0; k < kF; ++k) column_reduce(gw[k], red, out + k * kH);
  column_reduce(gb, red, out + kF * kH);
}

// -------------------------------------------------------------- update
// One thread per column p of the slot partials: the parameters p < count - 2,
// then the loss and the target energy. Mode 0 packs the current weights;
// mode 1 also sums the slot partials into grad, and loss[0..1] = (mean
// squared error, mean squared normalised target); mode 2 then takes an Adam
// step (beta 0.9, 0.999, epsilon 1e-8, bias-corrected at `step`) before
// packing.
__global__ void __launch_bounds__(kBlock) k_learned_update(int count, int mode, int slots,
    const float* __restrict__ partials, float* __restrict__ theta, float* __restrict__ am,
    float* __restrict__ av, float* __restrict__ grad, double* __restrict__ loss, int layers,
    int step, float lr, int nodes, bf16_t* __restrict__ wf, bf16_t* __restrict__ wb) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= count) return;
  double g = 0;
  if (mode >= 1) {
    for (int s = 0; s < slots; ++s) g += double(partials[long(s) * count + p]);
    if (p >= count - 2) {
      loss[p - (count - 2)] = g / double(nodes);
      return;
    }
    grad[p] = float(g);
  }
  if (p >= count - 2) return;
  float th = theta[p];
  if (mode == 2) {
    const float gf = float(g);
    const float m1 = 0.9f * am[p] + 0.1f * gf;
    const float v1 = 0.999f * av[p] + 0.001f * gf * gf;
    am[p] = m1;
    av[p] = v1;
    const double mh = double(m1) / (1.0 - pow(0.9, double(step)));
    const double vh = double(v1) / (1.0 - pow(0.999, double(step)));
    th = float(double(th) - double(lr) * mh / (sqrt(vh) + 1e-8));
    theta[p] = th;
  }
  const int rel = p - kInputParams;
  if (rel < 0 || rel >= layers * kLayerParams) return;
  const int l = rel / kLayerParams, o = rel % kLayerParams;
  if (o >= 2 * kH * kH) return;
  const int k = o / kH, c = o % kH;  // row k of [Ws; Wn], column c
  const bf16_t b = to_bf16(th);
  wb[long(l) * 2 * kH * kH + o] = b;
  wf[long(l) * 2 * kH * kH + c * 2 * kH + k] = b;
}

// ------------------------------------------------------- solver transfer
// r_a = sum over the fine rows i aggregated (through every level between)
// into coarse node a of the fine residual. A group of `group` lanes (a power
// of two up to 64, sized by the host to the mean member count) per coarse
// node: lane l adds members l, l + group, ..., then the group sums its lanes
// in FP64 in a fixed butterfly order.
__global__ void __launch_bounds__(kBlock) k_learned_gather(int nc, int group, const float* __restrict__ fine,
    const int* __restrict__ moff, const int* __restrict__ mem, float* __restrict__ r) {
  const long gid = long(blockIdx.x) * kBlock + threadIdx.x;
  const int a = int(gid / group), lane = int(gid % group);
  if (a >= nc) return;  // whole groups: a group never straddles a wave
  double s = 0;
  for (int k = moff[a] + lane; k < moff[a + 1]; k += group) s += double(fine[mem[k]]);
  for (int off = group >> 1; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) r[a] = float(s);
}
// z_i += e_{map_i} for the fine rows that belong to a coarse node.
__global__ void __launch_bounds__(kBlock) k_learned_prolong(int n, const int* __restrict__ map,
    const float* __restrict__ e, float* __restrict__ z) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int a = map[i];
  if (a >= 0) z[i] += e[a];
}

#define LI(k) static_cast<const int*>(buf(a, k))
#define LF(k) static_cast<const float*>(buf(a, k))
#define LW(k) static_cast<float*>(buf(a, k))
#define LD(k) static_cast<double*>(buf(a, k))
#define LCH(k) static_cast<const bf16_t*>(buf(a, k))
#define LH(k) static_cast<bf16_t*>(buf(a, k))

unsigned tiles_for(int n) { return unsigned((n + kTile - 1) / kTile); }
void need_nodes(int n, const char* what) {
  if (n < 1 || long(n) * kH > 0x7fffffffL) throw std::runtime_error(std::string(what) + ": bad node count");
}

// row_ptr, col, val, node_features (2n), ws, wt (nnz each), gstats (1),
// scratch (kLearnedSlots doubles)
void l_graph(kh_device* d, int n, const kh_arg* a) {
  need_nodes(n, "learned_graph");
  double* partials = LD(7);
  k_learned_graph_diag<<<dim3(kSlots), dim3(kBlock), 0, d->stream>>>(n, LI(0), LI(1), LF(2), LW(3), partials);
  k_learned_graph_weights<<<dim3(blocks_for(n)), dim3(kBlock), 0, d->stream>>>(n, LI(0), LI(1), LF(2), LF(3),
                                                                                LW(4), LW(5));
  k_learned_graph_scale<<<dim3(blocks_for(n)), dim3(kBlock), 0, d->stream>>>(n, LW(3), partials, LW(6));
}
// r, scratch (kLearnedSlots doubles)
void l_norm(kh_device* d, int n, const kh_arg* a) {
  need_nodes(n, "learned_norm");
  k_learned_norm<<<dim3(kSlots), dim3(kBlock), 0, d->stream>>>(n, LF(0), LD(1));
}
// r, norm scratch, node_features, gstats, theta, x, h0, stats
void l_input(kh_device* d, int n, const kh_arg* a) {
  need_nodes(n, "learned_input");
  k_learned_input<<<dim3(blocks_for(std::size_t(n) * 8)), dim3(kBlock), 0, d->stream>>>(
      n, LF(0), static_cast<const double*>(buf(a, 1)), LF(2), LF(3), LF(4), LW(5), LH(6), LW(7));
}
// row_ptr, col, ws, h_prev, wf, bias, h_out, m_out
void l_layer(kh_device* d, int n, const kh_arg* a) {
  need_nodes(n, "learned_layer");
  k_learned_layer<<<dim3(tiles_for(n)), dim3(kBlock), 0, d->stream>>>(n, LI(0), LI(1), LF(2), LCH(3),
                                                                      LCH(4), LF(5), LH(6), LH(7));
}
// h_L, head (wo, bo), stats, e
void l_head(kh_device* d, int n, const kh_arg* a) {
  need_nodes(n, "learned_head");
  k_learned_head<<<dim3(blocks_for(n)), dim3(kBlock), 0, d->stream>>>(n, LCH(0), LF(1), LF(2), LW(3));
}
// h_L, head, stats, target, y, dy, partials (at the wo slot), width
void l_head_loss(kh_device* d, int n, const kh_arg* a) {
  need_nodes(n, "learned_head_loss");
  k_learned_head_loss<<<dim3(kSlots), dim3(kBlock), 0, d->stream>>>(n, LCH(0), LF(1), LF(2), LF(3), LW(4),
                                                                    LW(5), LW(6), i32(a, 7));
}
// last, dy, head, p_in, dm_in, row_ptr, col, wt, h_l, wb, dz, p_out, dm_out,
// partials (at the bias slot), width
void l_back(kh_device* d, int n, const kh_arg* a) {
  need_nodes(n, "learned_back");
  k_learned_back<<<dim3(kSlots), dim3(kBlock), 0, d->stream>>>(n, i32(a, 0), LF(1), LF(2), LF(3), LF(4),
      LI(5), LI(6), LF(7), LCH(8), LCH(9), LH(10), LW(11), LW(12), LW(13), i32(a, 14));
}
// h_prev, m, dz, partials (at the Ws slot), width
void l_weight_grad(kh_device* d, int n, const kh_arg* a) {
  need_nodes(n, "learned_weight_grad");
  k_learned_weight_grad<<<dim3(kSlots), dim3(kBlock), 0, d->stream>>>(n, LCH(0), LCH(1), LCH(2), LW(3),
                                                                      i32(a, 4));
}
// p_in, dm_in, row_ptr, col, wt, h0, x, partials, width
void l_back_input(kh_device* d, int n, const kh_arg* a) {
  need_nodes(n, "learned_back_input");
  k_learned_back_input<<<dim3(kSlots), dim3(kBlock), 0, d->stream>>>(n, LF(0), LF(1), LI(2), LI(3), LF(4),
                                                                     LCH(5), LF(6), LW(7), i32(a, 8));
}
// n = parameters + 2; mode, partials, theta, adam m, adam v, grad, loss (2),
// layers, step, lr, nodes, wf, wb
void l_update(kh_device* d, int n, const kh_arg* a) {
  const int layers = i32(a, 7);
  if (layers < 1 || layers > KH_LEARNED_MAX_LAYERS ||
      n != kInputParams + layers * kLayerParams + kH + 1 + 2)
    throw std::runtime_error("learned_update: parameter count does not match the layer count");
  k_learned_update<<<dim3(blocks_for(n)), dim3(kBlock), 0, d->stream>>>(n, i32(a, 0), kSlots, LF(1), LW(2),
      LW(3), LW(4), LW(5), LD(6), layers, i32(a, 8), f32(a, 9), i32(a, 10), LH(11), LH(12));
}
// Fine residual b - A z, one thread per fine row.
__global__ void __launch_bounds__(kBlock) k_learned_residual(int n, const int* __restrict__ rp,
    const int* __restrict__ col, const float* __restrict__ val, const float* __restrict__ b,
    const float* __restrict__ z, float* __restrict__ out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float ri = b[i];
  for (int j = rp[i]; j < rp[i + 1]; ++j) ri -= val[j] * z[col[j]];
  out[i] = ri;
}
// n = coarse nodes; row_ptr, col, val (fine), b, z, member_offsets, members,
// r, fine scratch, fine rows, group
void l_restrict(kh_device* d, int n, const kh_arg* a) {
  const int fine = i32(a, 9), group = i32(a, 10);
  if (group < 1 || group > 64 || (group & (group - 1)) != 0)
    throw std::runtime_error("learned_restrict: the group must be a power of two up to 64");
  if (fine < 1) throw std::runtime_error("learned_restrict: bad fine row count");
  k_learned_residual<<<dim3(blocks_for(fine)), dim3(kBlock), 0, d->stream>>>(fine, LI(0), LI(1), LF(2), LF(3),
                                                                             LF(4), LW(8));
  k_learned_gather<<<dim3(blocks_for(std::size_t(n) * group)), dim3(kBlock), 0, d->stream>>>(
      n, group, LF(8), LI(5), LI(6), LW(7));
}
// n = fine rows; map, e, z
void l_prolong(kh_device* d, int n, const kh_arg* a) {
  k_learned_prolong<<<dim3(blocks_for(n)), dim3(kBlock), 0, d->stream>>>(n, LI(0), LF(1), LW(2));
}

}  // namespace

const std::vector<KernelEntry>& learned_kernels() {
  static const std::vector<KernelEntry> table = {
      {KH_LEARNED_GRAPH, "learned_graph", "bbbbbbbb", l_graph},
      {KH_LEARNED_NORM, "learned_norm", "bb", l_norm},
      {KH_LEARNED_INPUT, "learned_input", "bbbbbbbb", l_input},
      {KH_LEARNED_LAYER, "learned_layer", "bbbbbbbb", l_layer},
      {KH_LEARNED_HEAD, "learned_head", "bbbb", l_head},
      {KH_LEARNED_HEAD_LOSS, "learned_head_loss", "bbbbbbbi", l_head_loss},
      {KH_LEARNED_BACK, "learned_back", "ibbbbbbbbbbbbbi", l_back},
      {KH_LEARNED_WEIGHT_GRAD, "learned_weight_grad", "bbbbi", l_weight_grad},
      {KH_LEARNED_BACK_INPUT, "learned_back_input", "bbbbbbbbi", l_back_input},
      {KH_LEARNED_UPDATE, "learned_update", "ibbbbbbiifibb", l_update},
      {KH_LEARNED_RESTRICT, "learned_restrict", "bbbbbbbbbii", l_restrict},
      {KH_LEARNED_PROLONG, "learned_prolong", "bbb", l_prolong},
  };
  return table;
}

}  // namespace kh
