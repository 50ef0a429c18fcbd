// Interface load-covariance measures (deep_frame/topology_problem.py:573-684) on device-resident
// unit-wrench solutions: the unit flexibility F_w = W^T U (all pairs of load and solution columns)
// and the combined displacement fields U M (load cases, eigen-directions) whose element energies
// give the gradients.  FP64 throughout; HBM-bound streams over the k columns (the Gram on the FP64 matrix
// cores for up to 48 x 48 outputs).
#include <mutex>
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

// The same partial sums on the FP64 matrix cores, for ka, kb <= 48: the output is cut into 16 x 16 tiles
// (TI x TJ of them) and every wave accumulates all tiles over its 16-DOF blocks with
// v_mfma_f64_16x16x4f64.  In a 4-DOF step lane l supplies A[i][k] = a[16 ti + l % 16][d_k] and
// B[k][j] = b[16 tj + l % 16][d_k] with k = l / 16; a lane reads four consecutive DOFs of its columns once
// and feeds them to four steps, so the DOFs of a block are visited in a permuted but common order, which a
// sum over DOFs does not see.  Each lane then holds C[l / 16 + 4 v][l % 16], v = 0..3, of every tile.
// The four waves of a workgroup are added in a fixed order through LDS, so results are deterministic.
typedef double gram_d4 __attribute__((ext_vector_type(4)));

template <int TI, int TJ, typename T>
__global__ __launch_bounds__(256) void k_gram_mfma(const T* __restrict__ a, const T* __restrict__ b, long long n,
                                                   int ka, int kb, double* __restrict__ partial) {
  __shared__ double red[TI * TJ * 4][64];
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int c16 = l & 15, kq = l >> 4;
  long long oa[TI], ob[TJ];  // this lane's operand columns (clamped: extra columns only feed unused outputs)
#pragma unroll
  for (int t = 0; t < TI; ++t) oa[t] = (long long)min(16 * t + c16, ka - 1) * n;
#pragma unroll
  for (int t = 0; t < TJ; ++t) ob[t] = (long long)min(16 * t + c16, kb - 1) * n;
  gram_d4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = gram_d4{0.0, 0.0, 0.0, 0.0};
  const long long blocks16 = (n + 15) / 16;
  for (long long blk = (long long)blockIdx.x * 4 + wave; blk < blocks16; blk += (long long)gridDim.x * 4) {
    const long long d0 = blk * 16 + 4 * kq;
    double av[TI][4], bv[TJ][4];
    if (blk * 16 + 16 <= n) {
#pragma unroll
      for (int t = 0; t < TI; ++t)
#pragma unroll
        for (int s = 0; s < 4; ++s) av[t][s] = (double)a[oa[t] + d0 + s];
#pragma unroll
      for (int t = 0; t < TJ; ++t)
#pragma unroll
        for (int s = 0; s < 4; ++s) bv[t][s] = (double)b[ob[t] + d0 + s];
    } else {  // last, partial block: DOFs past n contribute zeros
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const bool in = d0 + s < n;
#pragma unroll
        for (int t = 0; t < TI; ++t) av[t][s] = in ? (double)a[oa[t] + d0 + s] : 0.0;
#pragma unroll
        for (int t = 0; t < TJ; ++t) bv[t][s] = in ? (double)b[ob[t] + d0 + s] : 0.0;
      }
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i][s], bv[j][s], acc[i][j], 0, 0, 0);
  }
  // waves 1, 2, 3 are added into wave 0's tiles in that order
  for (int w = 1; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) red[(i * TJ + j) * 4 + v][l] = acc[i][j][v];
    }
    __syncthreads();
    if (wave == 0) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[i][j][v] += red[(i * TJ + j) * 4 + v][l];
    }
    __syncthreads();
  }
  if (wave != 0) return;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * i + kq + 4 * v, c = 16 * j + c16;
        if (r < ka && c < kb) partial[(long long)(r * kb + c) * gridDim.x + blockIdx.x] = acc[i][j][v];
      }
}

template <int TI, typename T>
static void launch_gram_mfma(int tj, int blocks, hipStream_t s, const T* a, const T* b, long long n, int ka, int kb,
                             double* work) {
  switch (tj) {
    case 1: hipLaunchKernelGGL((k_gram_mfma<TI, 1, T>), dim3(blocks), dim3(256), 0, s, a, b, n, ka, kb, work); break;
    case 2: hipLaunchKernelGGL((k_gram_mfma<TI, 2, T>), dim3(blocks), dim3(256), 0, s, a, b, n, ka, kb, work); break;
    default: hipLaunchKernelGGL((k_gram_mfma<TI, 3, T>), dim3(blocks), dim3(256), 0, s, a, b, n, ka, kb, work); break;
  }
}

template <typename T>
static void gram_mfma(const T* a, const T* b, long long n, int ka, int kb, double* work, int blocks, hipStream_t s) {
  const int ti = (ka + 15) / 16, tj = (kb + 15) / 16;
  if (ti == 1)
    launch_gram_mfma<1, T>(tj, blocks, s, a, b, n, ka, kb, work);
  else if (ti == 2)
    launch_gram_mfma<2, T>(tj, blocks, s, a, b, n, ka, kb, work);
  else
    launch_gram_mfma<3, T>(tj, blocks, s, a, b, n, ka, kb, work);
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

// Y[m0 + j][n] (+)= sum_k U[k][n] M[k][m0 + j] on the FP64 matrix cores, j < 16 TJ: every wave takes 16-DOF
// blocks and forms the transposed tile Y^T = M^T U^T: in a step over four k lane l supplies
// A[i][kk] = M[k0 + kk][m0 + 16 t + i] and B[kk][d] = U[k0 + kk][d0 + d] with i = d = l % 16, kk = l / 16, so
// each entry of U is read by exactly one lane once; afterwards lane l holds output column 16 t + l / 16 + 4 v
// at DOF d0 + l % 16, so the sixteen lanes of a column store one contiguous run.
template <int TJ, typename T>
__global__ __launch_bounds__(256) void k_combine_mfma(const T* __restrict__ u, long long n, int k,
                                                      const double* __restrict__ mat, int mcols, int m0,
                                                      T* __restrict__ y, int accumulate) {
  __shared__ double Ms[GRAM_MAXK][16 * TJ];
  for (int v = threadIdx.x; v < GRAM_MAXK * 16 * TJ; v += blockDim.x) {
    const int r = v / (16 * TJ), c = v - r * (16 * TJ);
    Ms[r][c] = (r < k && m0 + c < mcols) ? mat[(long long)r * mcols + m0 + c] : 0.0;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int c16 = l & 15, kq = l >> 4;
  const long long blocks16 = (n + 15) / 16;
  for (long long blk = (long long)blockIdx.x * 4 + wave; blk < blocks16; blk += (long long)gridDim.x * 4) {
    const long long d0 = blk * 16;
    const bool in = d0 + c16 < n;
    gram_d4 acc[TJ];
#pragma unroll
    for (int t = 0; t < TJ; ++t) acc[t] = gram_d4{0.0, 0.0, 0.0, 0.0};
    double av[GRAM_MAXK / 4];  // all of this lane's U entries of the block, requested before the products
#pragma unroll
    for (int q = 0; q < GRAM_MAXK / 4; ++q) {
      const int kk = 4 * q + kq;
      av[q] = (in && kk < k) ? (double)u[(long long)kk * n + d0 + c16] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < GRAM_MAXK / 4; ++q) {
      if (4 * q >= k) break;
#pragma unroll
      for (int t = 0; t < TJ; ++t)
        acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(Ms[4 * q + kq][16 * t + c16], av[q], acc[t], 0, 0, 0);
    }
    if (!in) continue;
#pragma unroll
    for (int t = 0; t < TJ; ++t)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int col = m0 + 16 * t + kq + 4 * v;
        if (col >= mcols) continue;
        T* out = y + (long long)col * n + d0 + c16;
        *out = accumulate ? (T)((double)*out + acc[t][v]) : (T)acc[t][v];
      }
  }
}

// matrix cores: up to 64 output columns per pass over U (16-DOF blocks, four waves per workgroup)
template <typename T>
static void combine_mfma(const T* u, long long n, int k, const double* m, int mcols, T* y, int accumulate,
                         hipStream_t s) {
  const int mblocks = (int)std::min<long long>((n + 63) / 64, 4096);
  for (int m0 = 0; m0 < mcols; m0 += 64) {
    const int tj = std::min(4, (mcols - m0 + 15) / 16);
    switch (tj) {
      case 1: hipLaunchKernelGGL((k_combine_mfma<1, T>), dim3(mblocks), dim3(256), 0, s, u, n, k, m, mcols, m0, y, accumulate); break;
      case 2: hipLaunchKernelGGL((k_combine_mfma<2, T>), dim3(mblocks), dim3(256), 0, s, u, n, k, m, mcols, m0, y, accumulate); break;
      case 3: hipLaunchKernelGGL((k_combine_mfma<3, T>), dim3(mblocks), dim3(256), 0, s, u, n, k, m, mcols, m0, y, accumulate); break;
      default: hipLaunchKernelGGL((k_combine_mfma<4, T>), dim3(mblocks), dim3(256), 0, s, u, n, k, m, mcols, m0, y, accumulate); break;
    }
    DFMI_HIP(hipGetLastError());
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
  if (ka <= 48 && kb <= 48) {  // matrix cores: 16 x 16 output tiles, 64 DOFs per workgroup pass
    gram_mfma<double>(a, b, n, ka, kb, work, blocks, s);
  } else {
    hipLaunchKernelGGL(k_gram, dim3(blocks), dim3(GRAM_NT), 0, s, a, b, n, ka, kb, work);
  }
  DFMI_HIP(hipGetLastError());
  const int count = ka * kb;
  hipLaunchKernelGGL(k_sum_blocks, dim3(count), dim3(256), 0, s, work, blocks, count, scale, out);
  DFMI_HIP(hipGetLastError());
}

void combine_device(const double* u, long long n, int k, const double* m, int mcols, double* y, int accumulate,
                    hipStream_t s) {
  if (n < 1 || k < 1 || k > GRAM_MAXK || mcols < 1) throw std::invalid_argument("combine: need n >= 1, 1 <= k <= 64");
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  if (k < 32 || mcols <= 16) {  // few input or output columns: the vector kernel streams U and Y faster
    for (int m0 = 0; m0 < mcols; m0 += COMB_M) {
      hipLaunchKernelGGL(k_combine, dim3(blocks), dim3(256), 0, s, u, n, k, m, mcols, m0, y, accumulate);
      DFMI_HIP(hipGetLastError());
    }
    return;
  }
  combine_mfma<double>(u, n, k, m, mcols, y, accumulate, s);
}

// ---- allocation-free Gram -------------------------------------------------------------------------------
// The per-workgroup partial sums of dfmi_gram_static live in this file-scope device array (8 MiB: 256
// workgroups x 64 x 64 doubles), one copy per GPU for the life of the process.  Host code never passes the
// array's name to a kernel: its address on the current device comes from hipGetSymbolAddress, looked up once
// per device and cached.  Every workgroup writes all of its partials before the fixed-order block sum reads
// them, so the array needs no clearing; calls that overlap on one GPU share it and race.
constexpr int STATIC_GRAM_BLOCKS = 256;
constexpr int STATIC_GRAM_DEVICES = 64;
static __device__ __align__(256) double scratch[(size_t)STATIC_GRAM_BLOCKS * GRAM_MAXK * GRAM_MAXK];

static double* scratch_address() {
  static std::mutex guard;
  static double* addresses[STATIC_GRAM_DEVICES] = {};
  int device = 0;
  DFMI_HIP(hipGetDevice(&device));
  if (device < 0 || device >= STATIC_GRAM_DEVICES) throw std::runtime_error("device index beyond the scratch table");
  std::lock_guard<std::mutex> hold(guard);
  if (!addresses[device]) {
    void* address = nullptr;
    DFMI_HIP(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)));
    if (!address) throw std::runtime_error("hipGetSymbolAddress returned no address for the Gram scratch");
    addresses[device] = static_cast<double*>(address);
  }
  return addresses[device];
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

int dfmi_gram_static(const double* a_dev, const double* b_dev, int64_t n, int32_t ka, int32_t kb, double* out_dev,
                     void* hip_stream) {
  DFMI_GUARD3({
    if (!a_dev || !b_dev || !out_dev) throw std::invalid_argument("gram_static: null operand");
    if (n < 1 || ka < 1 || kb < 1 || ka > GRAM_MAXK || kb > GRAM_MAXK)
      throw std::invalid_argument("gram_static: need n >= 1 and 1 <= ka, kb <= 64 (the scratch holds 256 x 64 x 64 "
                                  "partial sums)");
    gram_device(a_dev, b_dev, (long long)n, ka, kb, out_dev, scratch_address(), STATIC_GRAM_BLOCKS, 1.0,
                (hipStream_t)hip_stream);
  })
}

int dfmi_combine(const double* u_dev, int64_t n, int32_t k, const double* m_dev, int32_t mcols, double* y_dev,
                 int32_t accumulate, void* hip_stream) {
  DFMI_GUARD3({ combine_device(u_dev, (long long)n, k, m_dev, mcols, y_dev, accumulate, (hipStream_t)hip_stream); })
}

}  // extern "C"
