// Rigid-body kernel of floating fixtures (deep_frame/topology_multigrid.py:29-38, 151-195, 224-246,
// 262-283): the six rigid modes about the domain centre, combined by the caller's coefficients into the
// nk modes the Dirichlet DOFs leave free, masked per level and orthonormalised.
//   * fine level: nk device vectors over the padded grid (column-major like the PCG vectors); the
//     projection x -= B (B^T x) removes them from right-hand sides, true residuals and V-cycle outputs;
//   * coarsest level: the kernel over the live DOFs (host, at most 40000 x 6), nk pinned DOFs chosen by
//     column-pivoted QR of its transpose, and the projected inverse G = P A_keep^-1 P with P = I - K K^T
//     formed once at set-up, so the coarsest solve stays one dense product.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "dfmi_internal.h"

namespace dfmi {

struct RigidArgs {
  double coef[6][6];  // coef[m][a]: weight of rigid mode m in kernel vector a
  double h[3];        // level spacing
  double center[3];   // rotation centre
};

// rigid mode m, component c, at offset (x, y, z) from the centre
__host__ __device__ inline double rigid_mode(int m, int c, double x, double y, double z) {
  switch (m) {
    case 0: return c == 0 ? 1.0 : 0.0;
    case 1: return c == 1 ? 1.0 : 0.0;
    case 2: return c == 2 ? 1.0 : 0.0;
    case 3: return c == 1 ? -z : (c == 2 ? y : 0.0);
    case 4: return c == 0 ? z : (c == 2 ? -x : 0.0);
    default: return c == 0 ? -y : (c == 1 ? x : 0.0);
  }
}

__global__ void k_rigid_vectors(Geom g, RigidArgs ra, int nk, const uint8_t* __restrict__ mask,
                                double* __restrict__ out) {
  const long long node = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long N = g.nodes();
  if (node >= N) return;
  const int l = (int)(node % g.nz), j = (int)((node / g.nz) % g.ny), i = (int)(node / ((long long)g.nz * g.ny));
  const double x = i * ra.h[0] - ra.center[0], y = j * ra.h[1] - ra.center[1], z = l * ra.h[2] - ra.center[2];
  const uint8_t mk = mask[node];
  for (int a = 0; a < nk; ++a)
    for (int c = 0; c < 3; ++c) {
      double v = 0.0;
      for (int m = 0; m < 6; ++m) v += ra.coef[m][a] * rigid_mode(m, c, x, y, z);
      out[(long long)a * N * 3 + node * 3 + c] = ((mk >> c) & 1) ? v : 0.0;
    }
}

// lower Cholesky factor of the nk x nk matrix g, inverted: returns L^-1 (row-major)
static std::vector<double> cholesky_inverse(const std::vector<double>& g, int nk) {
  std::vector<double> L(nk * nk, 0.0), inv(nk * nk, 0.0);
  double big = 0.0;
  for (int a = 0; a < nk; ++a) big = std::max(big, g[a * nk + a]);
  for (int i = 0; i < nk; ++i)
    for (int j = 0; j <= i; ++j) {
      double s = g[i * nk + j];
      for (int q = 0; q < j; ++q) s -= L[i * nk + q] * L[j * nk + q];
      if (i == j) {
        if (!(s > 1e-14 * big)) throw std::runtime_error("rigid kernel is rank deficient on the active DOFs");
        L[i * nk + i] = std::sqrt(s);
      } else {
        L[i * nk + j] = s / L[j * nk + j];
      }
    }
  for (int c = 0; c < nk; ++c)  // solve L x = e_c
    for (int i = 0; i < nk; ++i) {
      double s = i == c ? 1.0 : 0.0;
      for (int q = 0; q < i; ++q) s -= L[i * nk + q] * inv[q * nk + c];
      inv[i * nk + c] = s / L[i * nk + i];
    }
  return inv;
}

void fine_kernel(const Geom& g, const double h[3], const double center[3], const double* coef, int nk,
                 const uint8_t* mask, double* out, double* tmp, double* work, int work_blocks, double* small,
                 hipStream_t s) {
  if (nk < 1 || nk > 6) throw std::invalid_argument("rigid kernel takes 1 to 6 modes");
  RigidArgs ra{};
  for (int m = 0; m < 6; ++m)
    for (int a = 0; a < nk; ++a) ra.coef[m][a] = coef[m * nk + a];
  for (int d = 0; d < 3; ++d) ra.h[d] = h[d], ra.center[d] = center[d];
  const long long n = g.nodes() * 3;
  hipLaunchKernelGGL(k_rigid_vectors, dim3((unsigned)((g.nodes() + 255) / 256)), dim3(256), 0, s, g, ra, nk, mask,
                     out);
  DFMI_HIP(hipGetLastError());
  // Cholesky QR, twice: with B B^T = L L^T, the rows of L^-1 B are orthonormal
  for (int pass = 0; pass < 2; ++pass) {
    gram_device(out, out, n, nk, nk, small, work, work_blocks, 1.0, s);
    std::vector<double> g2(nk * nk);
    DFMI_HIP(hipMemcpyAsync(g2.data(), small, sizeof(double) * nk * nk, hipMemcpyDeviceToHost, s));
    DFMI_HIP(hipStreamSynchronize(s));
    const std::vector<double> inv = cholesky_inverse(g2, nk);
    std::vector<double> mt(nk * nk);  // combine weights m[r][a] = (L^-1)[a][r]
    for (int r = 0; r < nk; ++r)
      for (int a = 0; a < nk; ++a) mt[r * nk + a] = inv[a * nk + r];
    DFMI_HIP(hipMemcpyAsync(small, mt.data(), sizeof(double) * nk * nk, hipMemcpyHostToDevice, s));
    combine_device(out, n, nk, small, nk, tmp, 0, s);
    DFMI_HIP(hipMemcpyAsync(out, tmp, sizeof(double) * (size_t)nk * n, hipMemcpyDeviceToDevice, s));
    DFMI_HIP(hipStreamSynchronize(s));
  }
}

void coarsest_kernel(const Geom& g, const double h[3], const double center[3], const double* coef, int nk,
                     const std::vector<int>& live, std::vector<double>& kernel, std::vector<int>& pins) {
  const int n = (int)live.size();
  kernel.assign((size_t)n * nk, 0.0);
  for (int idx = 0; idx < n; ++idx) {
    const long long d = live[idx];
    const long long node = d / 3;
    const int c = (int)(d % 3);
    const int l = (int)(node % g.nz), j = (int)((node / g.nz) % g.ny), i = (int)(node / ((long long)g.nz * g.ny));
    const double x = i * h[0] - center[0], y = j * h[1] - center[1], z = l * h[2] - center[2];
    for (int a = 0; a < nk; ++a) {
      double v = 0.0;
      for (int m = 0; m < 6; ++m) v += coef[m * nk + a] * rigid_mode(m, c, x, y, z);
      kernel[(size_t)idx * nk + a] = v;
    }
  }
  // orthonormal columns: modified Gram-Schmidt, twice
  for (int pass = 0; pass < 2; ++pass)
    for (int a = 0; a < nk; ++a) {
      for (int b = 0; b < a; ++b) {
        double dot = 0.0;
        for (int idx = 0; idx < n; ++idx) dot += kernel[(size_t)idx * nk + a] * kernel[(size_t)idx * nk + b];
        for (int idx = 0; idx < n; ++idx) kernel[(size_t)idx * nk + a] -= dot * kernel[(size_t)idx * nk + b];
      }
      double norm = 0.0;
      for (int idx = 0; idx < n; ++idx) norm += kernel[(size_t)idx * nk + a] * kernel[(size_t)idx * nk + a];
      norm = std::sqrt(norm);
      if (!(norm > 1e-12)) throw std::runtime_error("rigid kernel is rank deficient on the coarsest level");
      for (int idx = 0; idx < n; ++idx) kernel[(size_t)idx * nk + a] /= norm;
    }
  // pins: column-pivoted QR of K^T (columns = DOFs): repeatedly the DOF whose remaining row of K is
  // largest, then that direction is removed from every row
  std::vector<double> rem(kernel);
  pins.clear();
  for (int t = 0; t < nk; ++t) {
    int best = -1;
    double bn = -1.0;
    for (int idx = 0; idx < n; ++idx) {
      double s = 0.0;
      for (int a = 0; a < nk; ++a) s += rem[(size_t)idx * nk + a] * rem[(size_t)idx * nk + a];
      if (s > bn) bn = s, best = idx;
    }
    if (best < 0 || !(bn > 0)) throw std::runtime_error("rigid kernel pins: no independent DOF left");
    pins.push_back(best);
    std::vector<double> q(nk);
    const double inv = 1.0 / std::sqrt(bn);
    for (int a = 0; a < nk; ++a) q[a] = rem[(size_t)best * nk + a] * inv;
    for (int idx = 0; idx < n; ++idx) {
      double dot = 0.0;
      for (int a = 0; a < nk; ++a) dot += rem[(size_t)idx * nk + a] * q[a];
      for (int a = 0; a < nk; ++a) rem[(size_t)idx * nk + a] -= dot * q[a];
    }
  }
}

// rows and columns of the pinned DOFs -> identity rows/columns (before the inverse)
__global__ void k_pin_dense(double* a, int n, const int* __restrict__ pins, int np) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * np) return;
  const int p = pins[idx / n], j = (int)(idx % n);
  const double v = j == p ? 1.0 : 0.0;
  a[(long long)p * n + j] = v;
  a[(long long)j * n + p] = v;
}

// the pinned DOFs carry no solution: zero their diagonal entry of the inverse
__global__ void k_unpin_dense(double* a, int n, const int* __restrict__ pins, int np) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < np) a[(long long)pins[t] * n + pins[t]] = 0.0;
}

// W[i][a] = sum_j A[j][i] K[j][a] (= (A K)[i][a] for the symmetric A; column reads are coalesced)
__global__ void k_dense_kernel_product(const double* __restrict__ a, int n, const double* __restrict__ kern, int nk,
                                       double* __restrict__ w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int j = 0; j < n; ++j) {
    const double v = a[(long long)j * n + i];
    for (int q = 0; q < nk; ++q) acc[q] = fma(v, kern[(long long)j * nk + q], acc[q]);
  }
  for (int q = 0; q < nk; ++q) w[(long long)i * nk + q] = acc[q];
}

// A <- P A P = A - K W^T - W K^T + (K S) K^T with W = A K, S = K^T W
__global__ void k_dense_project(double* a, int n, const double* __restrict__ kern, const double* __restrict__ w,
                                const double* __restrict__ ks, int nk) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)n * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  double d = 0.0;
  for (int q = 0; q < nk; ++q)
    d += kern[(long long)i * nk + q] * w[(long long)j * nk + q] + w[(long long)i * nk + q] * kern[(long long)j * nk + q] -
         ks[(long long)i * nk + q] * kern[(long long)j * nk + q];
  a[idx] -= d;
}

void pin_dense(double* a, int n, const int* pins_dev, int np, hipStream_t s) {
  hipLaunchKernelGGL(k_pin_dense, dim3((unsigned)(((long long)n * np + 255) / 256)), dim3(256), 0, s, a, n, pins_dev,
                     np);
  DFMI_HIP(hipGetLastError());
}

void project_dense_inverse(double* a, int n, const int* pins_dev, int np, const std::vector<double>& kernel, int nk,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_unpin_dense, dim3(1), dim3(64), 0, s, a, n, pins_dev, np);
  DFMI_HIP(hipGetLastError());
  double *kd = nullptr, *wd = nullptr, *ksd = nullptr;
  const size_t bytes = sizeof(double) * (size_t)n * nk;
  DFMI_HIP(hipMallocAsync((void**)&kd, bytes, s));
  DFMI_HIP(hipMallocAsync((void**)&wd, bytes, s));
  DFMI_HIP(hipMallocAsync((void**)&ksd, bytes, s));
  DFMI_HIP(hipMemcpyAsync(kd, kernel.data(), bytes, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_dense_kernel_product, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, n, kd, nk, wd);
  DFMI_HIP(hipGetLastError());
  std::vector<double> w((size_t)n * nk);
  DFMI_HIP(hipMemcpyAsync(w.data(), wd, bytes, hipMemcpyDeviceToHost, s));
  DFMI_HIP(hipStreamSynchronize(s));
  std::vector<double> sm(nk * nk, 0.0), ks((size_t)n * nk, 0.0);
  for (int i = 0; i < n; ++i)
    for (int p = 0; p < nk; ++p)
      for (int q = 0; q < nk; ++q) sm[p * nk + q] += kernel[(size_t)i * nk + p] * w[(size_t)i * nk + q];
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < nk; ++q) {
      double v = 0.0;
      for (int p = 0; p < nk; ++p) v += kernel[(size_t)i * nk + p] * sm[p * nk + q];
      ks[(size_t)i * nk + q] = v;
    }
  DFMI_HIP(hipMemcpyAsync(ksd, ks.data(), bytes, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_dense_project, dim3((unsigned)(((long long)n * n + 255) / 256)), dim3(256), 0, s, a, n, kd, wd,
                     ksd, nk);
  DFMI_HIP(hipGetLastError());
  DFMI_HIP(hipFreeAsync(kd, s));
  DFMI_HIP(hipFreeAsync(wd, s));
  DFMI_HIP(hipFreeAsync(ksd, s));
  DFMI_HIP(hipStreamSynchronize(s));
}

}  // namespace dfmi
