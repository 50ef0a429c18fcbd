// Column-batched vector kernels for PCG and the smoother: per-column reductions
// (wavefront shuffle + per-block partials, then a fixed-order second pass, so
// results are deterministic), AXPY-type updates, precision conversion and the
// flat <-> padded-grid layout changes.  Column r of a multi-vector is the
// contiguous slice v[r * n .. (r + 1) * n).
#include <stdexcept>

#include "dfmi_internal.h"

namespace dfmi {

int reduce_blocks(long long n) {
  long long b = (n + 1023) / 1024;
  return (int)std::max(1LL, std::min(b, 512LL));
}

template <typename T>
__global__ __launch_bounds__(256) void k_column_dot(const T* __restrict__ a, const T* __restrict__ b, long long n,
                                                     double* partial, const int* active) {
  const int r = blockIdx.y;
  double acc = 0.0;
  if (!active || active[r]) {
    const T* A = a + (long long)r * n;
    const T* B = b + (long long)r * n;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
      acc += (double)A[i] * (double)B[i];
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[(long long)r * gridDim.x + blockIdx.x] = acc;
}

template <typename T>
void column_dot(const T* a, const T* b, long long n, int k, double* partial, int nblocks, const int* active,
                hipStream_t s) {
  hipLaunchKernelGGL(k_column_dot<T>, dim3(nblocks, k), dim3(256), 0, s, a, b, n, partial, active);
  DFMI_HIP(hipGetLastError());
}
template void column_dot<float>(const float*, const float*, long long, int, double*, int, const int*, hipStream_t);
template void column_dot<double>(const double*, const double*, long long, int, double*, int, const int*,
                                 hipStream_t);

template <typename T>
__global__ __launch_bounds__(256) void k_column_absmax(const T* __restrict__ a, long long n, double* partial,
                                                        const int* active) {
  const int r = blockIdx.y;
  double acc = 0.0;
  if (!active || active[r]) {
    const T* A = a + (long long)r * n;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
      acc = fmax(acc, fabs((double)A[i]));
  }
  acc = block_max(acc);
  if (threadIdx.x == 0) partial[(long long)r * gridDim.x + blockIdx.x] = acc;
}

template <typename T>
void column_absmax(const T* a, long long n, int k, double* partial, int nblocks, const int* active,
                   hipStream_t s) {
  hipLaunchKernelGGL(k_column_absmax<T>, dim3(nblocks, k), dim3(256), 0, s, a, n, partial, active);
  DFMI_HIP(hipGetLastError());
}
template void column_absmax<float>(const float*, long long, int, double*, int, const int*, hipStream_t);
template void column_absmax<double>(const double*, long long, int, double*, int, const int*, hipStream_t);

__global__ __launch_bounds__(256) void k_reduce_partials(const double* partial, int nblocks, double* out, int op) {
  const int r = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nblocks; i += 256) {
    const double v = partial[(long long)r * nblocks + i];
    acc = op == 0 ? acc + v : fmax(acc, v);
  }
  acc = op == 0 ? block_sum(acc) : block_max(acc);
  if (threadIdx.x == 0) out[r] = acc;
}

void reduce_partials(const double* partial, int nblocks, int k, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(256), 0, s, partial, nblocks, out, 0);
  DFMI_HIP(hipGetLastError());
}

void reduce_max(const double* partial, int nblocks, int k, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(256), 0, s, partial, nblocks, out, 1);
  DFMI_HIP(hipGetLastError());
}

// first Chebyshev step from a zero iterate: x = c * Dinv * b
// column-major: grid (blocks over 3 * nodes, columns); interleaved: one thread per (node, column), its three values
template <typename T>
__global__ void k_cheb_start_cols(const T* __restrict__ b, const T* __restrict__ dinv, double c, T* __restrict__ x,
                                  long long n3, const int* active) {
  const int r = blockIdx.y;
  if (active && !active[r]) return;
  const T cc = (T)c;
  const long long off = (long long)r * n3;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n3; i += (long long)gridDim.x * 256)
    x[off + i] = cc * dinv[i] * b[off + i];
}

template <typename T>
__global__ void k_cheb_start(const T* __restrict__ b, const T* __restrict__ dinv, double c, T* __restrict__ x,
                             int slots, int k, const int* active) {
  // interleaved: slot e = n * k + r holds the three values at 3e
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= slots) return;
  const int n = e / k, r = e - n * k;
  if (active && !active[r]) return;
  const T cc = (T)c;
  const long long o = 3LL * e;
#pragma unroll
  for (int q = 0; q < 3; ++q) x[o + q] = cc * dinv[3 * n + q] * b[o + q];
}

template <typename T>
void cheb_start(const T* b, const T* dinv, double c, T* x, long long nodes, int k, const int* active, VLay lay,
                hipStream_t s) {
  if (lay.rs != 3 || k == 1) {  // column-major (or a single column)
    const long long n3 = nodes * 3;
    const unsigned bx = (unsigned)std::min<long long>((n3 + 255) / 256, 4096);
    hipLaunchKernelGGL(k_cheb_start_cols<T>, dim3(bx, k), dim3(256), 0, s, b, dinv, c, x, n3, active);
  } else {
    if (nodes * 3 * (long long)k > (long long)INT32_MAX) throw std::runtime_error("cheb_start: level too large");
    hipLaunchKernelGGL(k_cheb_start<T>, dim3((unsigned)((nodes * k + 255) / 256)), dim3(256), 0, s, b, dinv, c, x,
                       (int)(nodes * k), k, active);
  }
  DFMI_HIP(hipGetLastError());
}
template void cheb_start<float>(const float*, const float*, double, float*, long long, int, const int*, VLay,
                                hipStream_t);
template void cheb_start<double>(const double*, const double*, double, double*, long long, int, const int*, VLay,
                                 hipStream_t);

// deterministic pseudo-random start vector in [-0.5, 0.5) on live DOFs (power iteration)
__global__ void k_power_init(const uint8_t* __restrict__ mask, double* __restrict__ v, long long nodes, unsigned seed) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nodes * 3) return;
  unsigned long long z = (unsigned long long)i * 0x9E3779B97F4A7C15ULL + seed;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
  v[i] = ((mask[i / 3] >> (i % 3)) & 1) ? u : 0.0;
}

void power_init(const uint8_t* mask, double* v, long long nodes, unsigned seed, hipStream_t s) {
  hipLaunchKernelGGL(k_power_init, dim3((unsigned)((nodes * 3 + 255) / 256)), dim3(256), 0, s, mask, v, nodes, seed);
  DFMI_HIP(hipGetLastError());
}

// v = Dinv y * inv_norm[0] (dinv == nullptr: v = y * inv_norm[0]); v may alias y
__global__ void k_dinv_scale(const double* y, const double* __restrict__ dinv, const double* inv_norm, double* v,
                             long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = (dinv ? dinv[i] : 1.0) * y[i] * (inv_norm ? inv_norm[0] : 1.0);
}

void dinv_scale(const double* y, const double* dinv, const double* inv_norm, double* v, long long nodes,
                hipStream_t s) {
  hipLaunchKernelGGL(k_dinv_scale, dim3((unsigned)((nodes * 3 + 255) / 256)), dim3(256), 0, s, y, dinv, inv_norm, v,
                     nodes * 3);
  DFMI_HIP(hipGetLastError());
}

// out[0] = 1 / sqrt(in[0]) (0 when in[0] is not positive)
__global__ void k_inv_sqrt(const double* in, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = in[0] > 0 ? 1.0 / sqrt(in[0]) : 0.0;
}

void inv_sqrt(const double* in, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_inv_sqrt, dim3(1), dim3(64), 0, s, in, out);
  DFMI_HIP(hipGetLastError());
}

// Polak-Ribiere: beta = (r.z_new - r.z_old) / rz_old, rz = r.z_new
__global__ void k_pcg_beta(const double* rzn, const double* rzc, double* rz, double* beta, int k, const int* active) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  if (active && !active[c]) {
    beta[c] = 0.0;
    return;
  }
  beta[c] = rz[c] != 0.0 ? (rzn[c] - rzc[c]) / rz[c] : 0.0;
  rz[c] = rzn[c];
}

void pcg_beta(const double* rzn, const double* rzc, double* rz, double* beta, int k, const int* active,
              hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_beta, dim3((k + 63) / 64), dim3(64), 0, s, rzn, rzc, rz, beta, k, active);
  DFMI_HIP(hipGetLastError());
}

// dst[c] = src[c] for columns with select[c] != 0
template <typename T>
__global__ void k_copy_columns(const T* __restrict__ src, T* __restrict__ dst, long long n, const int* select) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || (select && !select[c])) return;
  dst[(long long)c * n + i] = src[(long long)c * n + i];
}

template <typename T>
void copy_columns(const T* src, T* dst, long long n, int k, const int* select, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_columns<T>, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s, src, dst, n, select);
  DFMI_HIP(hipGetLastError());
}
template void copy_columns<float>(const float*, float*, long long, int, const int*, hipStream_t);
template void copy_columns<double>(const double*, double*, long long, int, const int*, hipStream_t);

template <typename T>
__global__ void k_fill_zero(T* x, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = T(0);
}

template <typename T>
void fill_zero(T* x, long long n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n);
  DFMI_HIP(hipGetLastError());
}
template void fill_zero<float>(float*, long long, hipStream_t);
template void fill_zero<double>(double*, long long, hipStream_t);

// v[c] *= f[c] for selected columns (stores 0 where f[c] == 0)
template <typename T>
__global__ void k_scale_columns(T* __restrict__ v, long long n, const double* f, const int* select) {
  const int c = blockIdx.y;
  if (select && !select[c]) return;
  const T w = f ? (T)f[c] : T(0);
  const long long off = (long long)c * n;
  if (w == T(0)) {  // a zero factor stores zeros without reading the column
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) v[off + i] = T(0);
    return;
  }
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) v[off + i] *= w;
}

template <typename T>
void scale_columns(T* v, long long n, int k, const double* f, const int* select, hipStream_t s) {
  const unsigned bx = (unsigned)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_scale_columns<T>, dim3(bx, k), dim3(256), 0, s, v, n, f, select);
  DFMI_HIP(hipGetLastError());
}
template void scale_columns<float>(float*, long long, int, const double*, const int*, hipStream_t);
template void scale_columns<double>(double*, long long, int, const double*, const int*, hipStream_t);

// per-column scalars: dst[c] = src[c] (mode 0) or dst[c] *= src[c]^2 (mode 1) for selected columns
__global__ void k_scalars(const double* src, double* dst, const int* select, int k, int mode) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k || (select && !select[c])) return;
  dst[c] = mode == 0 ? src[c] : dst[c] * src[c] * src[c];
}

void select_scalars(const double* src, double* dst, const int* select, int k, bool square_scale, hipStream_t s) {
  hipLaunchKernelGGL(k_scalars, dim3((k + 63) / 64), dim3(64), 0, s, src, dst, select, k, square_scale ? 1 : 0);
  DFMI_HIP(hipGetLastError());
}

// x += scale[c] * d for active columns (iterative-refinement correction, FP32 -> FP64)
__global__ void k_add_scaled(double* __restrict__ x, float* __restrict__ d, const double* scale, long long n,
                             const int* active, bool zero_d) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || (active && !active[c])) return;
  const long long e = (long long)c * n + i;
  x[e] += scale[c] * (double)d[e];
  if (zero_d) d[e] = 0.f;
}

void add_scaled(double* x, float* d, const double* scale, long long n, int k, const int* active, hipStream_t s,
                bool zero_d) {
  hipLaunchKernelGGL(k_add_scaled, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s, x, d, scale, n, active,
                     zero_d);
  DFMI_HIP(hipGetLastError());
}

// x += alpha p, r -= alpha Ap, partial |r|^2 (columns with alpha = 0 are left unchanged); with z != nullptr
// also the partial sums of r_new . z (the r . z_old term of the Polak-Ribiere beta) in partial[k * nblocks ...]
template <typename T>
__global__ __launch_bounds__(256) void k_pcg_update(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                                                     const T* __restrict__ ap, const double* alpha, long long n,
                                                     const int* active, double* partial, const T* __restrict__ z) {
  const int c = blockIdx.y;
  double acc = 0.0, acz = 0.0;
  if (!active || active[c]) {
    const T a = (T)alpha[c];
    const long long off = (long long)c * n;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
      x[off + i] += a * p[off + i];
      const T rv = r[off + i] - a * ap[off + i];
      r[off + i] = rv;
      acc += (double)rv * (double)rv;
      if (z) acz += (double)rv * (double)z[off + i];
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[(long long)c * gridDim.x + blockIdx.x] = acc;
  if (z) {
    acz = block_sum(acz);
    if (threadIdx.x == 0) partial[((long long)gridDim.y + c) * gridDim.x + blockIdx.x] = acz;
  }
}

template <typename T>
void pcg_update(T* x, T* r, const T* p, const T* ap, const double* alpha, long long n, int k, const int* active,
                double* partial_rr, int nblocks, hipStream_t s, const T* z) {
  hipLaunchKernelGGL(k_pcg_update<T>, dim3(nblocks, k), dim3(256), 0, s, x, r, p, ap, alpha, n, active, partial_rr,
                     z);
  DFMI_HIP(hipGetLastError());
}
template void pcg_update<float>(float*, float*, const float*, const float*, const double*, long long, int, const int*,
                                double*, int, hipStream_t, const float*);
template void pcg_update<double>(double*, double*, const double*, const double*, const double*, long long, int,
                                 const int*, double*, int, hipStream_t, const double*);

template <typename T>
__global__ void k_pcg_direction(T* __restrict__ p, const T* __restrict__ z, const double* beta, long long n,
                                const int* active) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || (active && !active[c])) return;
  const long long o = (long long)c * n + i;
  p[o] = z[o] + (T)beta[c] * p[o];
}

template <typename T>
void pcg_direction(T* p, const T* z, const double* beta, long long n, int k, const int* active, hipStream_t s) {
  hipLaunchKernelGGL(k_pcg_direction<T>, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s, p, z, beta, n,
                     active);
  DFMI_HIP(hipGetLastError());
}
template void pcg_direction<float>(float*, const float*, const double*, long long, int, const int*, hipStream_t);
template void pcg_direction<double>(double*, const double*, const double*, long long, int, const int*, hipStream_t);

__global__ void k_alpha(const double* rz, const double* pap, double* alpha, int k, const int* active) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= k) return;
  alpha[c] = (active && !active[c]) || pap[c] == 0.0 ? 0.0 : rz[c] / pap[c];
}

void pcg_scalars_alpha(const double* rz, const double* pap, double* alpha, int k, const int* active,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_alpha, dim3((k + 63) / 64), dim3(64), 0, s, rz, pap, alpha, k, active);
  DFMI_HIP(hipGetLastError());
}

template <typename T>
__global__ void k_to_work(const double* __restrict__ r, const double* scale, T* __restrict__ w, long long n,
                          const int* active) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || (active && !active[c])) return;
  const double sc = scale[c] > 1e-300 ? 1.0 / scale[c] : 1.0;
  w[(long long)c * n + i] = (T)(r[(long long)c * n + i] * sc);
}

template <typename T>
void to_work(const double* r, const double* scale, T* w, long long n, int k, const int* active, hipStream_t s) {
  hipLaunchKernelGGL(k_to_work<T>, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s, r, scale, w, n, active);
  DFMI_HIP(hipGetLastError());
}
template void to_work<float>(const double*, const double*, float*, long long, int, const int*, hipStream_t);
template void to_work<double>(const double*, const double*, double*, long long, int, const int*, hipStream_t);

template <typename T>
__global__ void k_from_work(const T* __restrict__ w, const double* scale, double* __restrict__ z, long long n,
                            const int* active) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || (active && !active[c])) return;
  const double sc = scale[c] > 1e-300 ? scale[c] : 1.0;
  z[(long long)c * n + i] = (double)w[(long long)c * n + i] * sc;
}

template <typename T>
void from_work(const T* w, const double* scale, double* z, long long n, int k, const int* active, hipStream_t s) {
  hipLaunchKernelGGL(k_from_work<T>, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, s, w, scale, z, n, active);
  DFMI_HIP(hipGetLastError());
}
template void from_work<float>(const float*, const double*, double*, long long, int, const int*, hipStream_t);
template void from_work<double>(const double*, const double*, double*, long long, int, const int*, hipStream_t);

// flat DOF vectors (reference numbering over the unpadded (ux+1, uy+1, uz+1) nodes) <->
// padded grid vectors over pad.nodes(); padding nodes are zero
__global__ void k_flat_to_grid(Geom pad, int vx, int vy, int vz, const double* __restrict__ flat,
                               double* __restrict__ grid) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  const long long N = pad.nodes();
  if (n >= N) return;
  const int l = (int)(n % pad.nz), j = (int)((n / pad.nz) % pad.ny), i = (int)(n / ((long long)pad.nz * pad.ny));
  double* G = grid + ((long long)r * N + n) * 3;
  if (i < vx && j < vy && l < vz) {
    const long long u = ((long long)i * vy + j) * vz + l;
    const double* F = flat + ((long long)r * vx * vy * vz + u) * 3;
    G[0] = F[0];
    G[1] = F[1];
    G[2] = F[2];
  } else {
    G[0] = G[1] = G[2] = 0.0;
  }
}

void flat_to_grid(const Geom& pad, int ux, int uy, int uz, const double* flat, double* grid, int k, hipStream_t s) {
  hipLaunchKernelGGL(k_flat_to_grid, dim3((unsigned)((pad.nodes() + 255) / 256), k), dim3(256), 0, s, pad, ux + 1,
                     uy + 1, uz + 1, flat, grid);
  DFMI_HIP(hipGetLastError());
}

__global__ void k_grid_to_flat(Geom pad, int vx, int vy, int vz, const double* __restrict__ grid,
                               double* __restrict__ flat) {
  const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  const long long V = (long long)vx * vy * vz;
  if (u >= V) return;
  const int l = (int)(u % vz), j = (int)((u / vz) % vy), i = (int)(u / ((long long)vz * vy));
  const long long n = ((long long)i * pad.ny + j) * pad.nz + l;
  const double* G = grid + ((long long)r * pad.nodes() + n) * 3;
  double* F = flat + ((long long)r * V + u) * 3;
  F[0] = G[0];
  F[1] = G[1];
  F[2] = G[2];
}

void grid_to_flat(const Geom& pad, int ux, int uy, int uz, const double* grid, double* flat, int k, hipStream_t s) {
  const long long V = (long long)(ux + 1) * (uy + 1) * (uz + 1);
  hipLaunchKernelGGL(k_grid_to_flat, dim3((unsigned)((V + 255) / 256), k), dim3(256), 0, s, pad, ux + 1, uy + 1,
                     uz + 1, grid, flat);
  DFMI_HIP(hipGetLastError());
}

__global__ void k_pad_moduli(Geom pad, int ux, int uy, int uz, const double* __restrict__ m,
                             double* __restrict__ m64, float* __restrict__ m32) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= pad.elems()) return;
  const int l = (int)(e % pad.ez), j = (int)((e / pad.ez) % pad.ey), i = (int)(e / ((long long)pad.ez * pad.ey));
  double v = 0.0;
  if (i < ux && j < uy && l < uz) v = m[((long long)i * uy + j) * uz + l];
  m64[e] = v;
  m32[e] = (float)v;
}

void pad_moduli(const Geom& pad, int ux, int uy, int uz, const double* moduli, double* m64, float* m32,
                hipStream_t s) {
  hipLaunchKernelGGL(k_pad_moduli, dim3((unsigned)((pad.elems() + 255) / 256)), dim3(256), 0, s, pad, ux, uy, uz,
                     moduli, m64, m32);
  DFMI_HIP(hipGetLastError());
}

__global__ void k_scale_vector(double* v, double a, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= a;
}

void scale_vector(double* v, double a, long long n, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_vector, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v, a, n);
  DFMI_HIP(hipGetLastError());
}

__global__ void k_mask_vector(const uint8_t* __restrict__ mask, double* __restrict__ v, long long nodes) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (i >= nodes * 3) return;
  if (!((mask[i / 3] >> (i % 3)) & 1)) v[(long long)r * nodes * 3 + i] = 0.0;
}

void mask_vector(const uint8_t* mask, double* v, long long nodes, int k, hipStream_t s) {
  hipLaunchKernelGGL(k_mask_vector, dim3((unsigned)((nodes * 3 + 255) / 256), k), dim3(256), 0, s, mask, v, nodes);
  DFMI_HIP(hipGetLastError());
}

// ---- Krylov coarse solve on interleaved level vectors [node][r][3] ---------------------------
// Per-column dot products of up to four vector pairs in one pass over the vectors.  A workgroup owns a
// contiguous range of nodes; its threads form slots of 3 kc lanes (one lane per stored value of the kc
// columns c0 .. c0 + kc - 1 of a node), slot q taking nodes n0 + q, n0 + q + slots, ...  Each lane keeps
// its own sum; a column's 3 x slots lane sums are then added in a fixed order, so the partials are
// deterministic.  partial[(i * k + r) * gridDim.x + block] holds pair i, column r.
constexpr int KDOT_NT = 256;
constexpr int KDOT_COLS = KDOT_NT / 3;  // columns per launch (one slot of lanes per node at least)

template <typename T, int NP>
__global__ __launch_bounds__(KDOT_NT) void k_level_dots(LevelPairs<T> p, long long nodes, int k, int c0, int kc,
                                                        const int* active, double* partial) {
  __shared__ double lane[NP][KDOT_NT];
  const int w = 3 * kc, slots = KDOT_NT / w;
  const int t = threadIdx.x, q = t / w, j = t - q * w;
  double acc[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) acc[i] = 0.0;
  if (q < slots && (!active || active[c0 + j / 3])) {
    const long long per = (nodes + gridDim.x - 1) / gridDim.x;
    const long long n0 = (long long)blockIdx.x * per, n1 = min(nodes, n0 + per);
    for (long long n = n0 + q; n < n1; n += slots) {
      const long long o = n * 3LL * k + 3LL * c0 + j;
#pragma unroll
      for (int i = 0; i < NP; ++i) acc[i] += (double)p.a[i][o] * (double)p.b[i][o];
    }
  }
#pragma unroll
  for (int i = 0; i < NP; ++i) lane[i][t] = acc[i];
  __syncthreads();
  for (int e = t; e < NP * kc; e += KDOT_NT) {
    const int i = e / kc, r = e - i * kc;
    double sum = 0.0;
    for (int qq = 0; qq < slots; ++qq)
      for (int c = 0; c < 3; ++c) sum += lane[i][qq * w + 3 * r + c];
    partial[((long long)i * k + c0 + r) * gridDim.x + blockIdx.x] = sum;
  }
}

int level_dot_blocks(long long nodes) { return (int)std::max(1LL, std::min(2048LL, (nodes + 127) / 128)); }

template <typename T>
void level_dots(const LevelPairs<T>& p, int np, long long nodes, int k, const int* active, double* partial,
                int nblocks, hipStream_t s) {
  for (int c0 = 0; c0 < k; c0 += KDOT_COLS) {
    const int kc = std::min(KDOT_COLS, k - c0);
    if (np == 2)
      hipLaunchKernelGGL((k_level_dots<T, 2>), dim3(nblocks), dim3(KDOT_NT), 0, s, p, nodes, k, c0, kc, active,
                         partial);
    else if (np == 4)
      hipLaunchKernelGGL((k_level_dots<T, 4>), dim3(nblocks), dim3(KDOT_NT), 0, s, p, nodes, k, c0, kc, active,
                         partial);
    else
      throw std::invalid_argument("level_dots: 2 or 4 vector pairs");
    DFMI_HIP(hipGetLastError());
  }
}
template void level_dots<float>(const LevelPairs<float>&, int, long long, int, const int*, double*, int, hipStream_t);
template void level_dots<double>(const LevelPairs<double>&, int, long long, int, const int*, double*, int,
                                 hipStream_t);

// ks rows (k doubles each): 0 rho0 = g.z0, 1 z0.A z0, 2 rho1 = r1.z1, 3 r1.z0, 4 z1.A z1, 5 z1.A z0,
// 6 alpha0, 7 coefficient of z1, 8 coefficient of z0 in the result.
__global__ void k_krylov_alpha(double* ks, int k, const int* active) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k) return;
  const double rho = ks[r], zaz = ks[k + r];
  const bool ok = (!active || active[r]) && rho > 0.0 && zaz > 0.0;
  ks[6 * k + r] = ok ? rho / zaz : 0.0;
  ks[7 * k + r] = 0.0;
  ks[8 * k + r] = ok ? rho / zaz : 0.0;
}

// second flexible-CG step: beta = (r1.z1 - r1.z0) / rho0, p1 = z1 + beta z0 with
// p1.A p1 = z1.A z1 + 2 beta z1.A z0 + beta^2 z0.A z0, alpha1 = r1.z1 / p1.A p1, and the solution
// alpha0 z0 + alpha1 p1 = alpha1 z1 + (alpha0 + alpha1 beta) z0
__global__ void k_krylov_coefficients(double* ks, int k, const int* active) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= k) return;
  const double rho0 = ks[r], zaz0 = ks[k + r], rho1 = ks[2 * k + r], rz0 = ks[3 * k + r];
  const double zaz1 = ks[4 * k + r], zaz10 = ks[5 * k + r], a0 = ks[6 * k + r];
  double cz = 0.0, cp = a0;
  if ((!active || active[r]) && a0 > 0.0) {
    const double beta = (rho1 - rz0) / rho0;
    const double pap = zaz1 + 2.0 * beta * zaz10 + beta * beta * zaz0;
    if (pap > 0.0 && rho1 > 0.0) {
      const double a1 = rho1 / pap;
      cz = a1;
      cp = a0 + a1 * beta;
    }
  }
  ks[7 * k + r] = cz;
  ks[8 * k + r] = cp;
}

void krylov_scalars(double* ks, int k, const int* active, int step, hipStream_t s) {
  if (step == 1)
    hipLaunchKernelGGL(k_krylov_alpha, dim3((k + 63) / 64), dim3(64), 0, s, ks, k, active);
  else
    hipLaunchKernelGGL(k_krylov_coefficients, dim3((k + 63) / 64), dim3(64), 0, s, ks, k, active);
  DFMI_HIP(hipGetLastError());
}

// y -= f[r] x over interleaved level vectors, one thread per (node, column): the node's three values
// (columns with f = 0 or inactive are left alone)
template <typename T>
__global__ void k_level_axpy(T* __restrict__ y, const T* __restrict__ x, const double* f, int slots, int k,
                             const int* active) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= slots) return;
  const int r = e % k;
  if (active && !active[r]) return;
  const T fr = (T)f[r];
  const long long o = 3LL * e;
#pragma unroll
  for (int c = 0; c < 3; ++c) y[o + c] -= fr * x[o + c];
}

// x = a[r] x + b[r] p over interleaved level vectors, one thread per (node, column) (x is only read where
// a[r] != 0)
template <typename T>
__global__ void k_level_combine(T* __restrict__ x, const T* __restrict__ p, const double* a, const double* b,
                                int slots, int k, const int* active) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= slots) return;
  const int r = e % k;
  if (active && !active[r]) return;
  const double ca = a[r];
  const T cb = (T)b[r];
  const long long o = 3LL * e;
#pragma unroll
  for (int c = 0; c < 3; ++c) x[o + c] = (ca != 0.0 ? (T)ca * x[o + c] : T(0)) + cb * p[o + c];
}

static int level_slots(long long nodes, int k) {
  if (nodes * (long long)k > (long long)INT32_MAX) throw std::runtime_error("level vectors too large");
  return (int)(nodes * k);
}

template <typename T>
void level_axpy(T* y, const T* x, const double* f, long long nodes, int k, const int* active, hipStream_t s) {
  const int n = level_slots(nodes, k);
  hipLaunchKernelGGL(k_level_axpy<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y, x, f, n, k, active);
  DFMI_HIP(hipGetLastError());
}
template void level_axpy<float>(float*, const float*, const double*, long long, int, const int*, hipStream_t);
template void level_axpy<double>(double*, const double*, const double*, long long, int, const int*, hipStream_t);

template <typename T>
void level_combine(T* x, const T* p, const double* a, const double* b, long long nodes, int k, const int* active,
                   hipStream_t s) {
  const int n = level_slots(nodes, k);
  hipLaunchKernelGGL(k_level_combine<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, p, a, b, n, k,
                     active);
  DFMI_HIP(hipGetLastError());
}
template void level_combine<float>(float*, const float*, const double*, const double*, long long, int, const int*,
                                   hipStream_t);
template void level_combine<double>(double*, const double*, const double*, const double*, long long, int,
                                    const int*, hipStream_t);

}  // namespace dfmi
