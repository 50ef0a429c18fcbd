// Interface load-covariance measures (deep_frame/topology_problem.py:573-684) on device-resident
// unit-wrench solutions: the unit flexibility F_w = W^T U (all pairs of load and solution columns)
// and the combined displacement fields U M (load cases, eigen-directions) whose element energies
// give the gradients.  FP64 throughout; HBM-bound streams over the k columns.
#include <stdexcept>

#include "dfmi_internal.h"
#include "dfmi.h"

namespace dfmi {

constexpr int GRAM_NT = 256;
constexpr int GRAM_CH = 64;     // DOFs staged per pass
constexpr int GRAM_MAXK = 64;   // columns of each operand
constexpr int GRAM_T = 4;       // each thread owns a 4 x 4 tile of output pairs
static_assert(GRAM_NT * GRAM_T * GRAM_T <= GRAM_MAXK * (GRAM_CH + 1), "lane sums must fit the staging buffer");

// Per-block partial sums of C[i][j] = sum_n A[i][n] B[j][n] over the block's DOF chunks.  The output is
// cut into 4 x 4 tiles and every tile gets 256 / tiles threads, which split the chunk's DOFs between them
// (so small Gram matrices keep all threads busy); the tile's lanes are summed through LDS in a fixed order
// at the end.  partial is [output element][block], so the block sums read contiguous runs.
__global__ __launch_bounds__(GRAM_NT) void k_gram(const double* __restrict__ a, const double* __restrict__ b,
                                                   long long n, int ka, int kb, double* __restrict__ partial) {
  __shared__ double As[GRAM_MAXK][GRAM_CH + 1];
  __shared__ double Bs[GRAM_MAXK][GRAM_CH + 1];
  const int t = threadIdx.x;
  const int tiles_j = (kb + GRAM_T - 1) / GRAM_T;
  const int tiles = ((ka + GRAM_T - 1) / GRAM_T) * tiles_j;
  const int lanes = max(1, GRAM_NT / tiles);
  const int tile = t / lanes, lane = t - (t / lanes) * lanes;
  const bool owner = tile < tiles;
  const int ti = owner ? tile / tiles_j : 0, tj = owner ? tile - (tile / tiles_j) * tiles_j : 0;
  int ra[GRAM_T], rb[GRAM_T];  // operand rows of this tile, clamped into the staged rows
#pragma unroll
  for (int x = 0; x < GRAM_T; ++x) {
    ra[x] = min(ti * GRAM_T + x, ka - 1);
    rb[x] = min(tj * GRAM_T + x, kb - 1);
  }
  double acc[GRAM_T][GRAM_T];
#pragma unroll
  for (int x = 0; x < GRAM_T; ++x)
#pragma unroll
    for (int y = 0; y < GRAM_T; ++y) acc[x][y] = 0.0;
  for (long long c0 = (long long)blockIdx.x * GRAM_CH; c0 < n; c0 += (long long)gridDim.x * GRAM_CH) {
    for (int v = t; v < ka * GRAM_CH; v += GRAM_NT) {
      const int r = v / GRAM_CH, c = v - (v / GRAM_CH) * GRAM_CH;
      const long long idx = c0 + c;
      As[r][c] = idx < n ? a[(long long)r * n + idx] : 0.0;
    }
    for (int v = t; v < kb * GRAM_CH; v += GRAM_NT) {
      const int r = v / GRAM_CH, c = v - (v / GRAM_CH) * GRAM_CH;
      const long long idx = c0 + c;
      Bs[r][c] = idx < n ? b[(long long)r * n + idx] : 0.0;
    }
    __syncthreads();
    if (owner) {
      for (int c = lane; c < GRAM_CH; c += lanes) {
        double av[GRAM_T], bv[GRAM_T];
#pragma unroll
        for (int x = 0; x < GRAM_T; ++x) {
          av[x] = As[ra[x]][c];
          bv[x] = Bs[rb[x]][c];
        }
#pragma unroll
        for (int x = 0; x < GRAM_T; ++x)
#pragma unroll
          for (int y = 0; y < GRAM_T; ++y) acc[x][y] = fma(av[x], bv[y], acc[x][y]);
      }
    }
    __syncthreads();
  }
  // lanes of a tile -> one partial per output element, through the (now free) staging buffer
  double* red = &As[0][0];  // [GRAM_NT][GRAM_T * GRAM_T] fits in As
#pragma unroll
  for (int x = 0; x < GRAM_T; ++x)
#pragma unroll
    for (int y = 0; y < GRAM_T; ++y) red[t * GRAM_T * GRAM_T + x * GRAM_T + y] = acc[x][y];
  __syncthreads();
  for (int q = t; q < tiles * GRAM_T * GRAM_T; q += GRAM_NT) {
    const int tt = q / (GRAM_T * GRAM_T), e = q - tt * (GRAM_T * GRAM_T);
    const int i = (tt / tiles_j) * GRAM_T + e / GRAM_T, j = (tt - (tt / tiles_j) * tiles_j) * GRAM_T + e % GRAM_T;
    if (i >= ka || j >= kb) continue;
    double sum = 0.0;
    for (int l = 0; l < lanes; ++l) sum += red[(tt * lanes + l) * GRAM_T * GRAM_T + e];
    partial[(long long)(i * kb + j) * gridDim.x + blockIdx.x] = sum;
  }
}

// out[e] = scale * (sum over blocks of partial[e][block]): one workgroup per element, fixed-order tree
__global__ __launch_bounds__(256) void k_sum_blocks(const double* __restrict__ partial, int nblocks, int count,
                                                   double scale, double* __restrict__ out) {
  __shared__ double red[256];
  const int e = blockIdx.x, t = threadIdx.x;
  if (e >= count) return;
  double s = 0.0;
  for (int blk = t; blk < nblocks; blk += 256) s += partial[(long long)e * nblocks + blk];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) out[e] = scale * red[0];
}

constexpr int COMB_M = 16;  // output columns per pass

// Y[m][n] (+)= sum_k U[k][n] M[k][m] for output columns m0 .. m0 + COMB_M
__global__ __launch_bounds__(256) void k_combine(const double* __restrict__ u, long long n, int k,
                                                  const double* __restrict__ mat, int mcols, int m0,
                                                  double* __restrict__ y, int accumulate) {
  __shared__ double Ms[GRAM_MAXK][COMB_M];
  for (int v = threadIdx.x; v < GRAM_MAXK * COMB_M; v += blockDim.x) {
    const int r = v / COMB_M, c = v - (v / COMB_M) * COMB_M;
    Ms[r][c] = (r < k && m0 + c < mcols) ? mat[(long long)r * mcols + m0 + c] : 0.0;
  }
  __syncthreads();
  const int mc = min(COMB_M, mcols - m0);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double acc[COMB_M];
#pragma unroll
    for (int c = 0; c < COMB_M; ++c) acc[c] = 0.0;
    for (int r = 0; r < k; ++r) {
      const double v = u[(long long)r * n + i];
#pragma unroll
      for (int c = 0; c < COMB_M; ++c) acc[c] = fma(v, Ms[r][c], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < COMB_M; ++c)
      if (c < mc) {
        double* out = y + (long long)(m0 + c) * n + i;
        *out = accumulate ? *out + acc[c] : acc[c];
      }
  }
}

void gram_device(const double* a, const double* b, long long n, int ka, int kb, double* out, double* work,
                 int work_blocks, double scale, hipStream_t s) {
  if (n < 1 || ka < 1 || kb < 1 || ka > GRAM_MAXK || kb > GRAM_MAXK)
    throw std::invalid_argument("gram: need n >= 1 and 1 <= ka, kb <= 64");
  if (((ka + GRAM_T - 1) / GRAM_T) * ((kb + GRAM_T - 1) / GRAM_T) > GRAM_NT)
    throw std::invalid_argument("gram: too many output tiles");
  if (work_blocks < 1) throw std::invalid_argument("gram: work_blocks must be >= 1");
  const long long chunks = (n + GRAM_CH - 1) / GRAM_CH;
  const int blocks = (int)std::min<long long>(chunks, work_blocks);
  hipLaunchKernelGGL(k_gram, dim3(blocks), dim3(GRAM_NT), 0, s, a, b, n, ka, kb, work);
  DFMI_HIP(hipGetLastError());
  const int count = ka * kb;
  hipLaunchKernelGGL(k_sum_blocks, dim3(count), dim3(256), 0, s, work, blocks, count, scale, out);
  DFMI_HIP(hipGetLastError());
}

void combine_device(const double* u, long long n, int k, const double* m, int mcols, double* y, int accumulate,
                    hipStream_t s) {
  if (n < 1 || k < 1 || k > GRAM_MAXK || mcols < 1) throw std::invalid_argument("combine: need n >= 1, 1 <= k <= 64");
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  for (int m0 = 0; m0 < mcols; m0 += COMB_M) {
    hipLaunchKernelGGL(k_combine, dim3(blocks), dim3(256), 0, s, u, n, k, m, mcols, m0, y, accumulate);
    DFMI_HIP(hipGetLastError());
  }
}

}  // namespace dfmi

using namespace dfmi;

#define DFMI_GUARD3(...)                 \
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

int dfmi_gram(const double* a_dev, const double* b_dev, int64_t n, int32_t ka, int32_t kb, double* out_dev,
              double* work_dev, int32_t work_blocks, void* hip_stream) {
  DFMI_GUARD3({ gram_device(a_dev, b_dev, (long long)n, ka, kb, out_dev, work_dev, work_blocks, 1.0,
                            (hipStream_t)hip_stream); })
}

int dfmi_combine(const double* u_dev, int64_t n, int32_t k, const double* m_dev, int32_t mcols, double* y_dev,
                 int32_t accumulate, void* hip_stream) {
  DFMI_GUARD3({ combine_device(u_dev, (long long)n, k, m_dev, mcols, y_dev, accumulate, (hipStream_t)hip_stream); })
}

}  // extern "C"
