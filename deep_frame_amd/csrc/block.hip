// Passes over the DOFs of the block PCG recurrence (block Krylov solve over the k right-hand sides).
//
// Every FP32 multi-vector is column-major, [column][n].  One iteration of the recurrence needs four passes:
//   * a Gram pass  G = V^T [W | R]  (FP32 matrix cores, v_mfma_f32_16x16x4f32);
//   * an update    D += V a,  R -= W a,  with the new column norms |R_j|^2 (vector units);
//   * a Gram pass  H = W^T Z;
//   * a combination V' = Z + V b.
// The k x k coefficient matrices a and b are small and are formed between the passes (host or one
// workgroup), so each pass streams its operands once.  Sums over DOFs: FP32 inside a wave's share of the
// DOFs, FP64 across waves and workgroups, in a fixed order (deterministic).
#include <stdexcept>

#include "dfmi_internal.h"

namespace dfmi {

typedef float bk_f4 __attribute__((ext_vector_type(4)));

// Per-workgroup partial sums of C = A^T [B1 | B2] (ka <= 16 TI, kb1 <= 16 TJ1, kb2 <= 16 TJ2; TJ2 = 0: no B2).
// Each wave sweeps a contiguous range of 16-DOF blocks.  In a block lane l reads the four consecutive DOFs
// d0 + 4 (l / 16) .. + 3 of its columns 16 t + l % 16 with one 16-byte buffer load per operand tile (columns
// clamped to the last one: extra columns only feed outputs that are not written; DOFs at or past n read zeros
// through the buffer's range check), and step s of the 16 x 16 x 4 product uses DOF d0 + 4 kk + s as its k
// index kk = l / 16, for A and B alike, so each DOF of the block enters exactly once.  Lane l then holds
// C[16 ti + 4 (l / 16) + v][16 tj + l % 16], v = 0..3.  The four waves are added in a fixed order through LDS.
template <int TI, int TJ1, int TJ2>
__global__ __launch_bounds__(256) void k_bgram(const float* __restrict__ a, int ka, const float* __restrict__ b1,
                                                int kb1, const float* __restrict__ b2, int kb2, long long n,
                                                double* __restrict__ partial) {
  constexpr int TJ = TJ1 + TJ2;
  __shared__ float red[TI * TJ * 4][64];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int c16 = l & 15, kq = l >> 4;
  const long long cb = n * (long long)sizeof(float);  // bytes of one column
  // one buffer per operand, over all its columns: lane offsets stay below 2 GiB (checked on the host)
  const __amdgpu_buffer_rsrc_t ra = buffer(a, cb * ka), rb1 = buffer(b1, cb * kb1),
                               rb2 = buffer(TJ2 ? b2 : b1, cb * (TJ2 ? kb2 : kb1));
  unsigned oa[TI], ob[TJ];
#pragma unroll
  for (int t = 0; t < TI; ++t) oa[t] = (unsigned)(min(16 * t + c16, ka - 1) * cb);
#pragma unroll
  for (int t = 0; t < TJ1; ++t) ob[t] = (unsigned)(min(16 * t + c16, kb1 - 1) * cb);
#pragma unroll
  for (int t = 0; t < TJ2; ++t) ob[TJ1 + t] = (unsigned)(min(16 * t + c16, kb2 - 1) * cb);
  bk_f4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = bk_f4{0.f, 0.f, 0.f, 0.f};
  const long long blocks16 = (n + 15) / 16;
  const long long waves = (long long)gridDim.x * 4, wid = (long long)blockIdx.x * 4 + wave;
  const long long per = (blocks16 + waves - 1) / waves;
  const long long b0 = wid * per, b1e = min(blocks16, b0 + per);
  // operands of block blk: this lane's four DOFs d0 + 4 kq .. + 3 of each of its columns (past the column's
  // end a 16-byte load would read the next column, so the last, partial block takes single loads below n)
  auto fetch = [&](long long blk, bk_f4 (&av)[TI], bk_f4 (&bv)[TJ]) {
    const long long d0 = blk * 16 + 4 * kq;
    const unsigned off = (unsigned)(d0 * (long long)sizeof(float));
    if (blk * 16 + 16 <= n) {
#pragma unroll
      for (int t = 0; t < TI; ++t) av[t] = __builtin_bit_cast(bk_f4, __builtin_amdgcn_raw_buffer_load_b128(ra, oa[t] + off, 0, 0));
#pragma unroll
      for (int t = 0; t < TJ1; ++t) bv[t] = __builtin_bit_cast(bk_f4, __builtin_amdgcn_raw_buffer_load_b128(rb1, ob[t] + off, 0, 0));
#pragma unroll
      for (int t = 0; t < TJ2; ++t)
        bv[TJ1 + t] = __builtin_bit_cast(bk_f4, __builtin_amdgcn_raw_buffer_load_b128(rb2, ob[TJ1 + t] + off, 0, 0));
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const bool in = d0 + q < n;
        const unsigned oq = off + 4u * q;
#pragma unroll
        for (int t = 0; t < TI; ++t) av[t][q] = in ? __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(ra, oa[t] + oq, 0, 0)) : 0.f;
#pragma unroll
        for (int t = 0; t < TJ1; ++t)
          bv[t][q] = in ? __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rb1, ob[t] + oq, 0, 0)) : 0.f;
#pragma unroll
        for (int t = 0; t < TJ2; ++t)
          bv[TJ1 + t][q] = in ? __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rb2, ob[TJ1 + t] + oq, 0, 0)) : 0.f;
      }
    }
  };
  // two blocks in flight per wave: the operands of blocks blk + 2 and blk + 3 are requested before the
  // products of blocks blk and blk + 1
  auto products = [&](const bk_f4 (&av)[TI], const bk_f4 (&bv)[TJ]) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i][q], bv[j][q], acc[i][j], 0, 0, 0);
  };
  bk_f4 a0[TI], b0v[TJ], a1[TI], b1v[TJ];
  if (b0 < b1e) fetch(b0, a0, b0v);
  if (b0 + 1 < b1e) fetch(b0 + 1, a1, b1v);
  for (long long blk = b0; blk < b1e; blk += 2) {
    bk_f4 n0[TI], n0b[TJ], n1[TI], n1b[TJ];
    if (blk + 2 < b1e) fetch(blk + 2, n0, n0b);
    if (blk + 3 < b1e) fetch(blk + 3, n1, n1b);
    products(a0, b0v);
    if (blk + 1 < b1e) products(a1, b1v);
#pragma unroll
    for (int t = 0; t < TI; ++t) a0[t] = n0[t], a1[t] = n1[t];
#pragma unroll
    for (int t = 0; t < TJ; ++t) b0v[t] = n0b[t], b1v[t] = n1b[t];
  }
  // waves 3, 2, 1 are added through LDS in that order, then wave 0 adds its own tiles
  for (int w = 3; w >= 1; --w) {
    if (wave == w) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            float& cell = red[(i * TJ + j) * 4 + v][l];
            cell = (w == 3 ? 0.f : cell) + acc[i][j][v];
          }
    }
    __syncthreads();
  }
  if (wave != 0) return;
  const int kbt = kb1 + kb2;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double sum = (double)(acc[i][j][v] + red[(i * TJ + j) * 4 + v][l]);
        const int r = 16 * i + 4 * kq + v;
        const bool second = j >= TJ1;
        const int cj = 16 * (second ? j - TJ1 : j) + c16;
        if (r < ka && cj < (second ? kb2 : kb1))
          partial[(long long)(r * kbt + (second ? kb1 + cj : cj)) * gridDim.x + blockIdx.x] = sum;
      }
}

constexpr int BK_MAXK = 48;

// Y[j] = X[j] + sgn sum_i U[i] M[i][j] for j < kout (M FP32, row i at m + 48 i, zero past kout), for one or two
// output sets (SETS) that share U: set 0 is (x0, y0, m0), set 1 is (x1, y1, m1).  NORM (one set): also
// per-workgroup partials of |Y[j]|^2.  A workgroup of NW waves takes chunks of 256 consecutive rows: it stages
// the chunk of every input column in LDS (one 16-byte buffer load per lane and column, all issued before the
// first use), then wave w forms OG output columns of set w / (NW / SETS) for four rows per lane, reading U from
// LDS and the wave-uniform coefficients through scalar loads.  Every HBM access is 1 KB contiguous per wave and
// each input entry is read from HBM once.  Only the last chunk of the grid can be partial; it is peeled off the
// loop and takes single accesses (rows at or past n read zeros, their stores are dropped by the buffer's range
// check), so the loop itself carries no tests around its loads.  Y may alias X of its set: an entry of X is
// read only by the lane that writes it.  U must not alias an output.
__device__ __forceinline__ void bk_load4(__amdgpu_buffer_rsrc_t r, unsigned off, bool whole, float (&v)[4]) {
  if (whole) {
    const auto w = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = __uint_as_float(w[q]);
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off + 4u * q, 0, 0));
  }
}
__device__ __forceinline__ void bk_store4(__amdgpu_buffer_rsrc_t r, unsigned off, bool whole, const float (&v)[4]) {
  if (whole) {
    __attribute__((ext_vector_type(4))) unsigned w = {__float_as_uint(v[0]), __float_as_uint(v[1]),
                                                      __float_as_uint(v[2]), __float_as_uint(v[3])};
    __builtin_amdgcn_raw_buffer_store_b128(w, r, off, 0, 0);
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[q]), r, off + 4u * q, 0, 0);
  }
}

// Four consecutive FP64 entries (32 bytes) per lane; the partial form takes single dwords, so rows at or past the
// buffer's end read zeros and their stores are dropped.
__device__ __forceinline__ void bk_load4d(__amdgpu_buffer_rsrc_t r, unsigned off, bool whole, double (&v)[4]) {
  unsigned w[8];
  if (whole) {
    const auto a = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    const auto b = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16u, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = a[q], w[4 + q] = b[q];
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) w[q] = __builtin_amdgcn_raw_buffer_load_b32(r, off + 4u * q, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) v[q] = __hiloint2double((int)w[2 * q + 1], (int)w[2 * q]);
}
__device__ __forceinline__ void bk_store4d(__amdgpu_buffer_rsrc_t r, unsigned off, bool whole, const double (&v)[4]) {
  unsigned w[8];
#pragma unroll
  for (int q = 0; q < 4; ++q) w[2 * q] = (unsigned)__double2loint(v[q]), w[2 * q + 1] = (unsigned)__double2hiint(v[q]);
  if (whole) {
    __attribute__((ext_vector_type(4))) unsigned a = {w[0], w[1], w[2], w[3]}, b = {w[4], w[5], w[6], w[7]};
    __builtin_amdgcn_raw_buffer_store_b128(a, r, off, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b128(b, r, off + 16u, 0, 0);
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) __builtin_amdgcn_raw_buffer_store_b32(w[q], r, off + 4u * q, 0, 0);
  }
}

// Flushed output columns of a one-set combination (FLUSH): an output column j with flag[j] != 0 adds
// scale[j] * Y[j] into the FP64 vector x (same row layout, column stride n) and stores zero as Y[j].
struct BkFlush {
  double* x;
  const double* scale;
  const int* flag;
};

template <bool B>
struct bk_tag {
  static constexpr bool value = B;
};

template <int NW, int OG, int SETS, bool NORM, bool FLUSH = false>
__global__ __launch_bounds__(64 * NW) void k_bcombine(const float* __restrict__ u, const float* x0, float* y0,
                                                       const float* __restrict__ m0, const float* x1, float* y1,
                                                       const float* __restrict__ m1, float sgn, int kin, int kout,
                                                       long long n, double* __restrict__ partial, BkFlush fl) {
  static_assert(!NORM || SETS == 1, "column norms are formed for one output set");
  static_assert(!FLUSH || (SETS == 1 && !NORM), "flushed columns are formed by a plain one-set pass");
  constexpr int WPS = NW / SETS;  // waves per output set
  __shared__ __align__(16) float Us[BK_MAXK][256];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int set = wave / WPS, j0 = (wave - set * WPS) * OG;
  const int jn = min(OG, kout - j0);  // wave-uniform; waves past kout only stage
  const float* x = set ? x1 : x0;
  float* y = set ? y1 : y0;
  const float* m = set ? m1 : m0;
  const long long cbytes = n * (long long)sizeof(float);
  float ss[NORM ? OG : 1];
#pragma unroll
  for (int o = 0; o < (NORM ? OG : 1); ++o) ss[o] = 0.f;
  auto chunk = [&](long long r0, auto whole_tag) {
    constexpr bool whole = decltype(whole_tag)::value;
    const unsigned off = (unsigned)((r0 + 4LL * l) * (long long)sizeof(float));
    {  // stage the chunk of every input column: all loads first, then the LDS stores
      constexpr int PER = (BK_MAXK + NW - 1) / NW;
      float v[PER][4];
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int i = wave + q * NW;
        if (i < kin) bk_load4(buffer(u + (long long)i * n, cbytes), off, whole, v[q]);
      }
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int i = wave + q * NW;
        if (i < kin)
          *reinterpret_cast<float4*>(&Us[i][4 * l]) = make_float4(sgn * v[q][0], sgn * v[q][1], sgn * v[q][2], sgn * v[q][3]);
      }
    }
    float acc[OG][4];  // the sums start from X, read before any store of this lane
#pragma unroll
    for (int o = 0; o < OG; ++o)
      bk_load4(buffer(x + (long long)max(0, min(j0 + o, kout - 1)) * n, cbytes), off, whole, acc[o]);
    __syncthreads();
    if (jn > 0) {
#pragma unroll 2
      for (int i = 0; i < kin; ++i) {
        const float4 uv = *reinterpret_cast<const float4*>(&Us[i][4 * l]);
        const float* mi = m + (long long)i * BK_MAXK + j0;
#pragma unroll
        for (int o = 0; o < OG; ++o) {
          const float c = mi[o];
          acc[o][0] = fmaf(c, uv.x, acc[o][0]);
          acc[o][1] = fmaf(c, uv.y, acc[o][1]);
          acc[o][2] = fmaf(c, uv.z, acc[o][2]);
          acc[o][3] = fmaf(c, uv.w, acc[o][3]);
        }
      }
    }
    __syncthreads();  // the next chunk overwrites Us
#pragma unroll
    for (int o = 0; o < OG; ++o) {  // fully unrolled with a wave-uniform guard (a break keeps a loop that
                                    // indexes acc dynamically)
      if (o < jn) {
        if constexpr (FLUSH) {
          if (fl.flag[j0 + o]) {  // wave-uniform: x += s Y, and Y restarts at zero
            const double sj = fl.scale[j0 + o];
            const __amdgpu_buffer_rsrc_t xr = buffer(fl.x + (long long)(j0 + o) * n, 2 * cbytes);
            double xv[4];
            bk_load4d(xr, 2u * off, whole, xv);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              xv[q] += sj * (double)acc[o][q];
              acc[o][q] = 0.f;
            }
            bk_store4d(xr, 2u * off, whole, xv);
          }
        }
        bk_store4(buffer(y + (long long)(j0 + o) * n, cbytes), off, whole, acc[o]);
        if (NORM) {  // rows past n were read as zeros
#pragma unroll
          for (int q = 0; q < 4; ++q) ss[o] = fmaf(acc[o][q], acc[o][q], ss[o]);
        }
      }
    }
  };
  // each workgroup sweeps one contiguous range of chunks, so that its accesses stay within a few pages of every
  // column (a grid-stride order would move every column to a new page at every step)
  const long long chunks = (n + 255) / 256, per = (chunks + gridDim.x - 1) / gridDim.x;
  const long long c0 = (long long)blockIdx.x * per, c1 = min(chunks, c0 + per);
  const long long cw = min(c1, n / 256);  // chunks of this range that are whole
  for (long long c = c0; c < cw; ++c) chunk(c * 256, bk_tag<true>{});
  if (c0 < c1 && cw < c1) chunk(cw * 256, bk_tag<false>{});  // the partial last chunk (empty ranges take none)
  if (NORM) {  // each wave owns its columns: lanes are summed in FP64, one partial per column and workgroup
#pragma unroll
    for (int o = 0; o < OG; ++o) {
      double v = (double)ss[o];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (l == 0 && o < jn) partial[(long long)(j0 + o) * gridDim.x + blockIdx.x] = v;
    }
  }
}

// out[i][j] = scale * sum_q S[i][q] H[q][j] (S, H k x k row-major FP64; out FP32 with row stride 48, zero past
// column k): one workgroup
__global__ __launch_bounds__(256) void k_bsmall(const double* __restrict__ S, const double* __restrict__ H, int k,
                                                 double scale, float* __restrict__ out) {
  for (int e = threadIdx.x; e < k * BK_MAXK; e += blockDim.x) {
    const int i = e / BK_MAXK, j = e - (e / BK_MAXK) * BK_MAXK;
    double acc = 0.0;
    if (j < k)
      for (int q = 0; q < k; ++q) acc = fma(S[i * k + q], H[q * k + j], acc);
    out[e] = (float)(scale * acc);
  }
}

int block_gram_blocks(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 63) / 64, 512)); }

// workgroups of a combination pass (chunks of 256 rows; also the stride of the norm partials)
int block_combine_blocks(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 4096)); }

template <int TI, int TJ1, int TJ2>
static void launch_bgram(int blocks, hipStream_t s, const float* a, int ka, const float* b1, int kb1, const float* b2,
                         int kb2, long long n, double* work) {
  hipLaunchKernelGGL((k_bgram<TI, TJ1, TJ2>), dim3(blocks), dim3(256), 0, s, a, ka, b1, kb1, b2, kb2, n, work);
}

// out (ka x (kb1 + kb2), row-major, FP64) = A^T [B1 | B2]; B2 may be null (kb2 = 0).  work holds
// (ka * (kb1 + kb2)) * block_gram_blocks(n) doubles.
void block_gram(const float* a, int ka, const float* b1, int kb1, const float* b2, int kb2, long long n, double* out,
                double* work, hipStream_t s) {
  if (n < 1 || ka < 1 || kb1 < 1 || ka > BK_MAXK || kb1 > BK_MAXK || kb2 < 0 || kb2 > BK_MAXK || (kb2 > 0 && !b2))
    throw std::invalid_argument("block_gram: need n >= 1, 1 <= ka, kb1 <= 48 and 0 <= kb2 <= 48");
  if ((long long)n * (long long)sizeof(float) * std::max(ka, std::max(kb1, kb2)) > (long long)INT32_MAX)
    throw std::invalid_argument("block_gram: operands above 2 GiB");
  const int blocks = block_gram_blocks(n);
  const int ti = (ka + 15) / 16, tj1 = (kb1 + 15) / 16, tj2 = (kb2 + 15) / 16;
  if (ti == 3 && tj1 == 3 && tj2 == 3)
    launch_bgram<3, 3, 3>(blocks, s, a, ka, b1, kb1, b2, kb2, n, work);
  else if (ti == 3 && tj1 == 3 && tj2 == 0)
    launch_bgram<3, 3, 0>(blocks, s, a, ka, b1, kb1, b2, kb2, n, work);
  else if (ti == 2 && tj1 == 2 && tj2 == 2)
    launch_bgram<2, 2, 2>(blocks, s, a, ka, b1, kb1, b2, kb2, n, work);
  else if (ti == 2 && tj1 == 2 && tj2 == 0)
    launch_bgram<2, 2, 0>(blocks, s, a, ka, b1, kb1, b2, kb2, n, work);
  else if (ti == 1 && tj1 == 1 && tj2 == 1)
    launch_bgram<1, 1, 1>(blocks, s, a, ka, b1, kb1, b2, kb2, n, work);
  else if (ti == 1 && tj1 == 1 && tj2 == 0)
    launch_bgram<1, 1, 0>(blocks, s, a, ka, b1, kb1, b2, kb2, n, work);
  else
    throw std::invalid_argument("block_gram: operand column counts must share a 16-column tile count");
  DFMI_HIP(hipGetLastError());
  reduce_partials(work, blocks, ka * (kb1 + kb2), out, s);
}

template <bool NORM>
static void launch_bcombine(const float* u, const float* x, float* y, const float* m, float sgn, int kin, int kout,
                            long long n, double* partial, hipStream_t s) {
  if ((long long)n * (long long)sizeof(float) > (long long)INT32_MAX)
    throw std::invalid_argument("block combination: columns above 2 GiB");
  const dim3 grid(block_combine_blocks(n));
  switch ((kout + 15) / 16) {  // 16 output columns per wave
    case 1:
      hipLaunchKernelGGL((k_bcombine<1, 16, 1, NORM>), grid, dim3(64), 0, s, u, x, y, m, x, y, m, sgn, kin, kout, n,
                         partial, BkFlush{});
      break;
    case 2:
      hipLaunchKernelGGL((k_bcombine<2, 16, 1, NORM>), grid, dim3(128), 0, s, u, x, y, m, x, y, m, sgn, kin, kout, n,
                         partial, BkFlush{});
      break;
    default:
      hipLaunchKernelGGL((k_bcombine<3, 16, 1, NORM>), grid, dim3(192), 0, s, u, x, y, m, x, y, m, sgn, kin, kout, n,
                         partial, BkFlush{});
      break;
  }
  DFMI_HIP(hipGetLastError());
}

void block_combine(const float* u, const float* x, float* y, const float* m, int kin, int kout, long long n,
                   hipStream_t s) {
  if (n < 1 || kin < 1 || kout < 1 || kin > BK_MAXK || kout > BK_MAXK)
    throw std::invalid_argument("block_combine: need n >= 1 and 1 <= kin, kout <= 48");
  launch_bcombine<false>(u, x, y, m, 1.f, kin, kout, n, nullptr, s);
}

// Y0 = X0 + U M0 and Y1 = X1 + U M1 in one pass over U (k columns in and out; M row stride 48): the next search
// block and the deferred solution update of the block recurrence.  8 output columns per wave, two sets of waves.
void block_combine2(const float* u, const float* x0, float* y0, const float* m0, const float* x1, float* y1,
                    const float* m1, int k, long long n, hipStream_t s) {
  if (n < 1 || k < 1 || k > BK_MAXK) throw std::invalid_argument("block_combine2: need n >= 1 and 1 <= k <= 48");
  if ((long long)n * (long long)sizeof(float) > (long long)INT32_MAX)
    throw std::invalid_argument("block combination: columns above 2 GiB");
  const dim3 grid(block_combine_blocks(n));
  switch ((k + 15) / 16) {  // waves per set: 2, 4 or 6 of 8 columns
    case 1:
      hipLaunchKernelGGL((k_bcombine<4, 8, 2, false>), grid, dim3(256), 0, s, u, x0, y0, m0, x1, y1, m1, 1.f, k, k, n,
                         nullptr, BkFlush{});
      break;
    case 2:
      hipLaunchKernelGGL((k_bcombine<8, 8, 2, false>), grid, dim3(512), 0, s, u, x0, y0, m0, x1, y1, m1, 1.f, k, k, n,
                         nullptr, BkFlush{});
      break;
    default:
      hipLaunchKernelGGL((k_bcombine<12, 8, 2, false>), grid, dim3(768), 0, s, u, x0, y0, m0, x1, y1, m1, 1.f, k, k,
                         n, nullptr, BkFlush{});
      break;
  }
  DFMI_HIP(hipGetLastError());
}

// D += U M (k columns, M row stride 48), and for every column j with flag[j] != 0 also x_j += scale[j] D_j (FP64,
// column stride n) followed by D_j = 0, in the same pass: the reliable update of the block recurrence, which moves
// the FP32 solution increment of the flushed columns into the FP64 solution.
void block_combine_flush(const float* u, float* d, const float* m, int k, long long n, double* x, const double* scale,
                         const int* flag, hipStream_t s) {
  if (n < 1 || k < 1 || k > BK_MAXK) throw std::invalid_argument("block_combine_flush: need n >= 1 and 1 <= k <= 48");
  if (!x || !scale || !flag) throw std::invalid_argument("block_combine_flush: need x, scale and flags");
  if ((long long)n * (long long)sizeof(double) > (long long)INT32_MAX)
    throw std::invalid_argument("block combination: FP64 columns above 2 GiB");
  const dim3 grid(block_combine_blocks(n));
  const BkFlush fl{x, scale, flag};
  switch ((k + 15) / 16) {
    case 1:
      hipLaunchKernelGGL((k_bcombine<1, 16, 1, false, true>), grid, dim3(64), 0, s, u, d, d, m, d, d, m, 1.f, k, k, n,
                         nullptr, fl);
      break;
    case 2:
      hipLaunchKernelGGL((k_bcombine<2, 16, 1, false, true>), grid, dim3(128), 0, s, u, d, d, m, d, d, m, 1.f, k, k,
                         n, nullptr, fl);
      break;
    default:
      hipLaunchKernelGGL((k_bcombine<3, 16, 1, false, true>), grid, dim3(192), 0, s, u, d, d, m, d, d, m, 1.f, k, k,
                         n, nullptr, fl);
      break;
  }
  DFMI_HIP(hipGetLastError());
}

// D += V M and R -= W M (k columns, M row stride 48), rr[j] = |R_j|^2 after the update; work holds
// k * block_combine_blocks(n) doubles.
void block_update(const float* v, float* d, const float* w, float* r, const float* m, int k, long long n, double* rr,
                  double* work, hipStream_t s) {
  if (n < 1 || k < 1 || k > BK_MAXK) throw std::invalid_argument("block_update: need n >= 1 and 1 <= k <= 48");
  launch_bcombine<false>(v, d, d, m, 1.f, k, k, n, nullptr, s);
  block_residual_update(w, r, m, k, n, rr, work, s);
}

// R -= W M alone, with rr[j] = |R_j|^2 (the solution update deferred to the next combination pass)
void block_residual_update(const float* w, float* r, const float* m, int k, long long n, double* rr, double* work,
                           hipStream_t s) {
  if (n < 1 || k < 1 || k > BK_MAXK) throw std::invalid_argument("block_update: need n >= 1 and 1 <= k <= 48");
  launch_bcombine<true>(w, r, r, m, -1.f, k, k, n, work, s);
  reduce_partials(work, block_combine_blocks(n), k, rr, s);
}

// out = scale * S H (k x k; out with row stride 48)
void block_small_product(const double* S, const double* H, int k, double scale, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_bsmall, dim3(1), dim3(256), 0, s, S, H, k, scale, out);
  DFMI_HIP(hipGetLastError());
}

}  // namespace dfmi

// ---- C ABI: the passes of the block recurrence ---------------------------------------------------------
#include "dfmi.h"

#define DFMI_BLOCK_GUARD(...)            \
  try {                                  \
    __VA_ARGS__;                         \
    return 0;                            \
  } catch (const std::exception& e) {    \
    dfmi::set_error(e.what());           \
    return 1;                            \
  } catch (...) {                        \
    dfmi::set_error("unknown failure");  \
    return 1;                            \
  }

extern "C" {

int dfmi_block_work_doubles(int64_t n, int32_t ka, int32_t kb, int64_t* count) {
  DFMI_BLOCK_GUARD({
    if (n < 1 || !count) throw std::invalid_argument("block_work_doubles: need n >= 1 and an output");
    const long long g = (long long)ka * kb * dfmi::block_gram_blocks(n);
    const long long c = (long long)dfmi::BK_MAXK * dfmi::block_combine_blocks(n);
    *count = g > c ? g : c;
  })
}

int dfmi_block_gram(const float* a_dev, int32_t ka, const float* b1_dev, int32_t kb1, const float* b2_dev,
                    int32_t kb2, int64_t n, double* out_dev, double* work_dev, void* hip_stream) {
  DFMI_BLOCK_GUARD({
    if (!a_dev || !b1_dev || !out_dev || !work_dev) throw std::invalid_argument("block_gram: null operand");
    dfmi::block_gram(a_dev, ka, b1_dev, kb1, b2_dev, kb2, (long long)n, out_dev, work_dev, (hipStream_t)hip_stream);
  })
}

int dfmi_block_combine(const float* u_dev, const float* x_dev, float* y_dev, const float* m_dev, int32_t kin,
                       int32_t kout, int64_t n, void* hip_stream) {
  DFMI_BLOCK_GUARD({
    if (!u_dev || !x_dev || !y_dev || !m_dev) throw std::invalid_argument("block_combine: null operand");
    if (u_dev == y_dev) throw std::invalid_argument("block_combine: y must not alias u");
    dfmi::block_combine(u_dev, x_dev, y_dev, m_dev, kin, kout, (long long)n, (hipStream_t)hip_stream);
  })
}

int dfmi_block_combine2(const float* u_dev, const float* x0_dev, float* y0_dev, const float* m0_dev,
                        const float* x1_dev, float* y1_dev, const float* m1_dev, int32_t k, int64_t n,
                        void* hip_stream) {
  DFMI_BLOCK_GUARD({
    if (!u_dev || !x0_dev || !y0_dev || !m0_dev || !x1_dev || !y1_dev || !m1_dev)
      throw std::invalid_argument("block_combine2: null operand");
    if (u_dev == y0_dev || u_dev == y1_dev) throw std::invalid_argument("block_combine2: y must not alias u");
    dfmi::block_combine2(u_dev, x0_dev, y0_dev, m0_dev, x1_dev, y1_dev, m1_dev, k, (long long)n,
                         (hipStream_t)hip_stream);
  })
}

int dfmi_block_combine_flush(const float* u_dev, float* d_dev, const float* m_dev, int32_t k, int64_t n,
                             double* x_dev, const double* scale_dev, const int32_t* flag_dev, void* hip_stream) {
  DFMI_BLOCK_GUARD({
    if (!u_dev || !d_dev || !m_dev) throw std::invalid_argument("block_combine_flush: null operand");
    if (u_dev == d_dev) throw std::invalid_argument("block_combine_flush: d must not alias u");
    dfmi::block_combine_flush(u_dev, d_dev, m_dev, k, (long long)n, x_dev, scale_dev, flag_dev,
                              (hipStream_t)hip_stream);
  })
}

int dfmi_block_update(const float* v_dev, float* d_dev, const float* w_dev, float* r_dev, const float* m_dev,
                      int32_t k, int64_t n, double* rr_dev, double* work_dev, void* hip_stream) {
  DFMI_BLOCK_GUARD({
    if (!v_dev || !d_dev || !w_dev || !r_dev || !m_dev || !rr_dev || !work_dev)
      throw std::invalid_argument("block_update: null operand");
    dfmi::block_update(v_dev, d_dev, w_dev, r_dev, m_dev, k, (long long)n, rr_dev, work_dev,
                       (hipStream_t)hip_stream);
  })
}

}  // extern "C"
