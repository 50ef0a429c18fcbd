// SIMP step around the solve: interpolation (deep_frame/topology_optimization.py:657-661),
// compliance/sensitivity reduction (:721-734), cone density filter and pullback
// (:1082-1092, :1135-1142, :1171-1178), OC update (:1197-1233) and one MMA step
// (Svanberg's method as called by deep_frame/topology_optimizers.py:55-67).
// All element-wise work is HBM-bound; reductions use per-block partials and a
// fixed-order second pass.
#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <vector>

#include "dfmi_internal.h"
#include "dfmi.h"

namespace dfmi {

static __device__ double blk_sum(double v) {
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

static __device__ double blk_max(double v) {
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// ---- SIMP interpolation ---------------------------------------------------------------------
__global__ void k_simp(const double* rho, const uint8_t* active, long long n, double young, double p, double eps,
                       double* moduli, double* dmoduli) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i];
  const bool on = !active || active[i];
  moduli[i] = on ? young * (eps + (1.0 - eps) * pow(r, p)) : 0.0;
  if (dmoduli) dmoduli[i] = on ? young * (1.0 - eps) * p * pow(r, p - 1.0) : 0.0;
}

// Heaviside projection of the filtered field (DensityMap.project, topology_optimization.py:1152-1171) at
// up to three thresholds in one pass: physical[f][i], derivative[f][i].  flags bit 0: preserve (physical
// 1), bit 1: forbidden (physical 0), bit 2: free (derivative kept).  The robust form writes
// sech^2 as 4 e^(-2|s|) / (1 + e^(-2|s|))^2, the single form as 1 - tanh^2, as the reference does.
struct ProjectArgs {
  double beta;
  double eta[3];
  int count;
  int robust;
};

__global__ void k_project(long long n, const double* __restrict__ filtered, const uint8_t* __restrict__ flags,
                          ProjectArgs pa, double* __restrict__ physical, double* __restrict__ derivative) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double f = filtered[i];
  const uint8_t fl = flags[i];
  const double b = pa.beta;
  for (int q = 0; q < pa.count; ++q) {
    const double eta = pa.eta[q];
    double phys, der;
    if (b != 0.0) {
      const double tb = tanh(b * eta);
      const double den = tb + tanh(b * (1.0 - eta));
      const double shifted = b * (f - eta);
      const double th = tanh(shifted);
      phys = (tb + th) / den;
      if (pa.robust) {
        const double decay = exp(-2.0 * fabs(shifted));
        der = b * 4.0 * decay / ((1.0 + decay) * (1.0 + decay)) / den;
      } else {
        der = b * (1.0 - th * th) / den;
      }
    } else {
      phys = f;
      der = 1.0;
    }
    if (fl & 1) phys = 1.0;
    if (fl & 2) phys = 0.0;
    if (!(fl & 4)) der = 0.0;
    physical[(long long)q * n + i] = fmin(fmax(phys, 0.0), 1.0);
    derivative[(long long)q * n + i] = der;
  }
}

// sensitivity[e] = -dmod[e] * factor * sum_r energy[r][e]
__global__ void k_sensitivity(const double* energy, long long ne, int k, const double* dmod, double factor,
                              double* sens) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double acc = 0.0;
  for (int r = 0; r < k; ++r) acc += energy[(long long)r * ne + e];
  sens[e] = -dmod[e] * factor * acc;
}

__global__ void k_flat_dot(const double* a, const double* b, long long n, double* partial) {
  const int r = blockIdx.y;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    acc += a[(long long)r * n + i] * b[(long long)r * n + i];
  acc = blk_sum(acc);
  if (threadIdx.x == 0) partial[(long long)r * gridDim.x + blockIdx.x] = acc;
}

__global__ void k_finish_compliance(const double* partial, int nb, double factor, double* out) {
  const int r = blockIdx.x;
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) acc += partial[(long long)r * nb + i];
  acc = blk_sum(acc);
  if (threadIdx.x == 0) out[r] = factor * acc;
}

void compliance_sensitivity(const double* u_flat, const double* f_flat, long long ndof, int k, const double* energy,
                            long long ne, const double* dmod, double factor, double* comp, double* sens,
                            double* partial, hipStream_t s) {
  const int nb = reduce_blocks(ndof);
  hipLaunchKernelGGL(k_flat_dot, dim3(nb, k), dim3(256), 0, s, u_flat, f_flat, ndof, partial);
  hipLaunchKernelGGL(k_finish_compliance, dim3(k), dim3(256), 0, s, partial, nb, factor, comp);
  hipLaunchKernelGGL(k_sensitivity, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, energy, ne, k, dmod, factor,
                     sens);
  DFMI_HIP(hipGetLastError());
}

// ---- cone filter -----------------------------------------------------------------------------
struct FilterGeom {
  int nx, ny, nz;
  int taps;
};

// out[e] = a_e sum_o w_o (x a)[e + o] * (scale ? scale[e] : 1), optionally zeroed outside `keep`
__global__ void k_cone(FilterGeom g, const int* offs, const double* w, const uint8_t* allowed, const double* x,
                       const double* pre, const double* post, const uint8_t* keep, double* out) {
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int l = (int)(e % g.nz), j = (int)((e / g.nz) % g.ny), i = (int)(e / ((long long)g.nz * g.ny));
  double acc = 0.0;
  if (allowed[e] && !(keep && !keep[e])) {
    for (int t = 0; t < g.taps; ++t) {
      const int a = i + offs[3 * t], b = j + offs[3 * t + 1], c = l + offs[3 * t + 2];
      if (a < 0 || b < 0 || c < 0 || a >= g.nx || b >= g.ny || c >= g.nz) continue;
      const long long m = ((long long)a * g.ny + b) * g.nz + c;
      if (!allowed[m]) continue;
      acc += w[t] * x[m] * (pre ? pre[m] : 1.0);
    }
  }
  if (post) acc *= post[e];
  if (keep && !keep[e]) acc = 0.0;
  out[e] = acc;
}

// corridor blend of the cone filter (topology_optimization.py:1125-1129, :1142, :1176):
// a = (1 - w) a + w b; lo = (1 - w) s d, hi = w s d; out += a
__global__ void k_blend(double* a, const double* b, const double* w, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = (1.0 - w[i]) * a[i] + w[i] * b[i];
}

__global__ void k_mul_split(const double* s, const double* d, const double* w, double* lo, double* hi, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = s[i] * d[i];
  lo[i] = (1.0 - w[i]) * v;
  hi[i] = w[i] * v;
}

__global__ void k_add(double* out, const double* a, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += a[i];
}

__global__ void k_inverse(const double* a, double* b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = 1.0 / a[i];
}

__global__ void k_sums_fix(const uint8_t* allowed, double* sums, long long n, int* bad) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!allowed[i]) sums[i] = 1.0;
  if (!(sums[i] > 0)) atomicAdd(bad, 1);
}

__global__ void k_mul(const double* a, const double* b, double* out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] * (b ? b[i] : 1.0);
}

__global__ void k_sum(const double* a, long long n, double* partial) {
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) acc += a[i];
  acc = blk_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// cone taps max(r - |o h|, 0) > 0 over |o_a| <= floor(r / h_a) (topology_optimization.py:1082-1092)
static void cone_taps(const double spacing[3], double radius, std::vector<int>& offs, std::vector<double>& w) {
  int pad[3];
  for (int a = 0; a < 3; ++a) pad[a] = (int)std::floor(radius / spacing[a]);
  for (int i = -pad[0]; i <= pad[0]; ++i)
    for (int j = -pad[1]; j <= pad[1]; ++j)
      for (int l = -pad[2]; l <= pad[2]; ++l) {
        const double dd =
            std::sqrt(std::pow(i * spacing[0], 2) + std::pow(j * spacing[1], 2) + std::pow(l * spacing[2], 2));
        const double wt = std::max(radius - dd, 0.0);
        if (wt > 0) {
          offs.insert(offs.end(), {i, j, l});
          w.push_back(wt);
        }
      }
}

struct Filter {
  FilterGeom g{};
  double spacing[3]{};
  hipStream_t stream = nullptr;
  long long n = 0;
  void *offs = nullptr, *w = nullptr, *allowed = nullptr, *free_ = nullptr, *sums = nullptr, *inv = nullptr,
       *tmp = nullptr, *tmp2 = nullptr, *partial = nullptr, *lo = nullptr, *hi = nullptr, *ratio = nullptr,
       *prop = nullptr;
  // corridor blend: the wide cone (g2 taps), blend weights w in [0, 1] on allowed cells, their support and
  // the wide cone's inverse sums (1 off the support)
  bool corridor = false;
  FilterGeom g2{};
  void *offs2 = nullptr, *w2 = nullptr, *blend = nullptr, *support = nullptr, *inv2 = nullptr, *tmp3 = nullptr;
  int nb = 1;
  ~Filter() {
    for (void* p : {offs, w, allowed, free_, sums, inv, tmp, tmp2, partial, lo, hi, ratio, prop, offs2, w2, blend,
                    support, inv2, tmp3})
      if (p) (void)hipFree(p);
  }
  double sum(const double* a) {
    hipLaunchKernelGGL(k_sum, dim3(nb), dim3(256), 0, stream, a, n, (double*)partial);
    std::vector<double> h(nb);
    DFMI_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * nb, hipMemcpyDeviceToHost, stream));
    DFMI_HIP(hipStreamSynchronize(stream));
    double acc = 0.0;
    for (double v : h) acc += v;
    return acc;
  }
  // filtered field x~ = (W (x a)) a / sums, blended with the wide cone on the corridor: (1 - w) x~ + w x~_wide
  void forward(const double* x, double* out) {
    const unsigned ge = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_cone, dim3(ge), dim3(256), 0, stream, g, (const int*)offs, (const double*)w,
                       (const uint8_t*)allowed, x, nullptr, (const double*)inv, nullptr, out);
    if (corridor) {
      hipLaunchKernelGGL(k_cone, dim3(ge), dim3(256), 0, stream, g2, (const int*)offs2, (const double*)w2,
                         (const uint8_t*)allowed, x, nullptr, (const double*)inv2, (const uint8_t*)support,
                         (double*)tmp3);
      hipLaunchKernelGGL(k_blend, dim3(ge), dim3(256), 0, stream, out, (const double*)tmp3, (const double*)blend, n);
    }
    DFMI_HIP(hipGetLastError());
  }
  // adjoint: a (W (a (1 - w) s d / sums) + W_wide (a w s d / sums_wide)), zero outside the free cells
  void pullback(const double* sens, const double* dproj, double* out) {
    const unsigned ge = (unsigned)((n + 255) / 256);
    if (!corridor) {
      hipLaunchKernelGGL(k_mul, dim3(ge), dim3(256), 0, stream, sens, dproj, (double*)tmp, n);
    } else {
      hipLaunchKernelGGL(k_mul_split, dim3(ge), dim3(256), 0, stream, sens, dproj, (const double*)blend, (double*)tmp,
                         (double*)tmp2, n);
    }
    hipLaunchKernelGGL(k_cone, dim3(ge), dim3(256), 0, stream, g, (const int*)offs, (const double*)w,
                       (const uint8_t*)allowed, (const double*)tmp, (const double*)inv, nullptr,
                       (const uint8_t*)free_, out);
    if (corridor) {
      hipLaunchKernelGGL(k_cone, dim3(ge), dim3(256), 0, stream, g2, (const int*)offs2, (const double*)w2,
                         (const uint8_t*)allowed, (const double*)tmp2, (const double*)inv2, nullptr,
                         (const uint8_t*)free_, (double*)tmp3);
      hipLaunchKernelGGL(k_add, dim3(ge), dim3(256), 0, stream, out, (const double*)tmp3, n);
    }
    DFMI_HIP(hipGetLastError());
  }
};

// ---- OC --------------------------------------------------------------------------------------
__global__ void k_oc_prep(const double* x, const double* dc, const double* dv, const uint8_t* free_, double dvmax,
                          double xmin, double move, double* ratio, double* lo, double* hi, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!free_[i]) return;
  const double d = fmax(dv[i], 1e-12 * dvmax);
  ratio[i] = fmax(1e-30, -dc[i] / d);
  lo[i] = fmax(xmin, x[i] - move);
  hi[i] = fmin(1.0, x[i] + move);
}

__global__ void k_oc_prop(const double* x, const uint8_t* free_, const double* ratio, const double* lo,
                          const double* hi, double lmb, int lowest, double* out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (!free_[i]) {
    out[i] = x[i];
    return;
  }
  out[i] = lowest ? lo[i] : fmin(fmax(x[i] * sqrt(ratio[i] / fmax(lmb, 1e-100)), lo[i]), hi[i]);
}

// min/max/finite check of dv and dc over free cells: partial[0..nb) min dv, [nb..2nb) max dv, [2nb..3nb) bad count
__global__ void k_oc_check(const double* dc, const double* dv, const uint8_t* free_, long long n, double* partial,
                           int nb) {
  double mn = 1e300, mx = -1e300, bad = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    if (!free_[i]) continue;
    mn = fmin(mn, dv[i]);
    mx = fmax(mx, dv[i]);
    if (!isfinite(dv[i]) || !isfinite(dc[i])) bad += 1.0;
  }
  const double a = -blk_max(-mn);
  const double b = blk_max(mx);
  const double c = blk_sum(bad);
  if (threadIdx.x == 0) {
    partial[blockIdx.x] = a;
    partial[nb + blockIdx.x] = b;
    partial[2 * nb + blockIdx.x] = c;
  }
}

__global__ void k_ratio_max(const double* ratio, const uint8_t* free_, long long n, double* partial) {
  double mx = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    if (free_[i]) mx = fmax(mx, ratio[i]);
  mx = blk_max(mx);
  if (threadIdx.x == 0) partial[blockIdx.x] = mx;
}

// ---- MMA (m constraints, a = 0) -------------------------------------------------------------------
// Svanberg's MMA as mmapy.mmasub runs it for deep_frame/topology_optimizers.py:55-67 (a0 = 1, a = 0, c = 1e4,
// d = 1, bounds [0, 1]): asymptote update, move-limited bounds, the convex separable approximation and the
// primal-dual interior-point subproblem.  The n-sized passes run here; the m x m Newton system of the dual
// variables is solved on the host (m <= MMA_MAX_M).
constexpr int MMA_MAX_M = 16;
constexpr int MMA_PAIRS = MMA_MAX_M * (MMA_MAX_M + 1) / 2;

struct MmaLam {
  double v[MMA_MAX_M];
};

__global__ void k_mma_prepare(long long n, int m, int it, const double* x, const double* xold1, const double* xold2,
                              double* low, double* upp, const double* df0, const double* dg, double move,
                              double scale, double* alfa, double* beta, double* p0, double* q0, double* pp,
                              double* qq, double* xs, double* xsi, double* eta, double* partial, int nb) {
  const double asyinit = 0.5, asydecr = 0.7, asyincr = 1.2, raa0 = 1e-5, albefa = 0.1;
  double acc[MMA_MAX_M];
  for (int j = 0; j < m; ++j) acc[j] = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double xv = x[i], span = 1.0;  // bounds [0, 1]
    double lo, up;
    if (it <= 2) {
      lo = xv - asyinit * span;
      up = xv + asyinit * span;
    } else {
      const double sg = (xv - xold1[i]) * (xold1[i] - xold2[i]);
      const double f = sg > 0 ? asyincr : (sg < 0 ? asydecr : 1.0);
      lo = xv - f * (xold1[i] - low[i]);
      up = xv + f * (upp[i] - xold1[i]);
      lo = fmin(fmax(lo, xv - 10.0 * span), xv - 0.01 * span);
      up = fmax(fmin(up, xv + 10.0 * span), xv + 0.01 * span);
    }
    low[i] = lo;
    upp[i] = up;
    const double al = fmax(fmax(lo + albefa * (xv - lo), xv - move * span), 0.0);
    const double be = fmin(fmin(up - albefa * (up - xv), xv + move * span), 1.0);
    alfa[i] = al;
    beta[i] = be;
    const double ux2 = (up - xv) * (up - xv), xl2 = (xv - lo) * (xv - lo);
    const double inv = 1.0 / fmax(span, 1e-5);
    const double d0 = scale * df0[i];
    const double a0p = fmax(d0, 0.0), a0q = fmax(-d0, 0.0);
    const double pq0 = 0.001 * (a0p + a0q) + raa0 * inv;
    p0[i] = (a0p + pq0) * ux2;
    q0[i] = (a0q + pq0) * xl2;
    for (int j = 0; j < m; ++j) {
      const double d1 = scale * dg[(long long)j * n + i];
      const double a1p = fmax(d1, 0.0), a1q = fmax(-d1, 0.0);
      const double pq = 0.001 * (a1p + a1q) + raa0 * inv;
      const double pv = (a1p + pq) * ux2, qv = (a1q + pq) * xl2;
      pp[(long long)j * n + i] = pv;
      qq[(long long)j * n + i] = qv;
      acc[j] += pv / (up - xv) + qv / (xv - lo);
    }
    const double x0 = 0.5 * (al + be);
    xs[i] = x0;
    xsi[i] = fmax(1.0 / (x0 - al), 1.0);
    eta[i] = fmax(1.0 / (be - x0), 1.0);
  }
  for (int j = 0; j < m; ++j) {
    const double t = blk_sum(acc[j]);
    if (threadIdx.x == 0) partial[(long long)j * nb + blockIdx.x] = t;
  }
}

struct MmaVecs {
  const double *low, *upp, *alfa, *beta, *p0, *q0, *pp, *qq;
};

// Newton system pieces: per constraint j the sums of gvec_j and of GG_j delx / diagx, and per pair (j, l),
// j <= l, the sum of GG_j GG_l / diagx (rows of partial: m, m, m (m + 1) / 2)
__global__ void k_mma_newton(long long n, int m, MmaVecs v, const double* x, const double* xsi, const double* eta,
                             MmaLam lam, double epsi, double* partial, int nb) {
  double sg[MMA_MAX_M], sgd[MMA_MAX_M], sgg[MMA_PAIRS];
  for (int j = 0; j < m; ++j) sg[j] = sgd[j] = 0.0;
  for (int q = 0; q < m * (m + 1) / 2; ++q) sgg[q] = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double ux = v.upp[i] - x[i], xl = x[i] - v.low[i];
    double plam = v.p0[i], qlam = v.q0[i];
    for (int j = 0; j < m; ++j) {
      plam += lam.v[j] * v.pp[(long long)j * n + i];
      qlam += lam.v[j] * v.qq[(long long)j * n + i];
    }
    const double delx = plam / (ux * ux) - qlam / (xl * xl) - epsi / (x[i] - v.alfa[i]) + epsi / (v.beta[i] - x[i]);
    const double diagx = 2.0 * (plam / (ux * ux * ux) + qlam / (xl * xl * xl)) + xsi[i] / (x[i] - v.alfa[i]) +
                         eta[i] / (v.beta[i] - x[i]);
    double gg[MMA_MAX_M];
    for (int j = 0; j < m; ++j) {
      const double pj = v.pp[(long long)j * n + i], qj = v.qq[(long long)j * n + i];
      gg[j] = pj / (ux * ux) - qj / (xl * xl);
      sg[j] += pj / ux + qj / xl;
      sgd[j] += gg[j] * delx / diagx;
    }
    int q = 0;
    for (int j = 0; j < m; ++j)
      for (int l = j; l < m; ++l) sgg[q++] += gg[j] * gg[l] / diagx;
  }
  for (int j = 0; j < m; ++j) {
    const double a = blk_sum(sg[j]);
    const double b = blk_sum(sgd[j]);
    if (threadIdx.x == 0) {
      partial[(long long)j * nb + blockIdx.x] = a;
      partial[(long long)(m + j) * nb + blockIdx.x] = b;
    }
  }
  for (int q = 0; q < m * (m + 1) / 2; ++q) {
    const double c = blk_sum(sgg[q]);
    if (threadIdx.x == 0) partial[(long long)(2 * m + q) * nb + blockIdx.x] = c;
  }
}

__global__ void k_mma_direction(long long n, int m, MmaVecs v, const double* x, const double* xsi, const double* eta,
                                MmaLam lam, MmaLam dlam, double epsi, double* dx, double* dxsi, double* deta,
                                double* partial) {
  double mx = -1e300;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double ux = v.upp[i] - x[i], xl = x[i] - v.low[i];
    double plam = v.p0[i], qlam = v.q0[i], ggd = 0.0;
    for (int j = 0; j < m; ++j) {
      const double pj = v.pp[(long long)j * n + i], qj = v.qq[(long long)j * n + i];
      plam += lam.v[j] * pj;
      qlam += lam.v[j] * qj;
      ggd += (pj / (ux * ux) - qj / (xl * xl)) * dlam.v[j];
    }
    const double xa = x[i] - v.alfa[i], bx = v.beta[i] - x[i];
    const double delx = plam / (ux * ux) - qlam / (xl * xl) - epsi / xa + epsi / bx;
    const double diagx = 2.0 * (plam / (ux * ux * ux) + qlam / (xl * xl * xl)) + xsi[i] / xa + eta[i] / bx;
    const double d = -delx / diagx - ggd / diagx;
    const double ds = -xsi[i] + epsi / xa - xsi[i] * d / xa;
    const double de = -eta[i] + epsi / bx + eta[i] * d / bx;
    dx[i] = d;
    dxsi[i] = ds;
    deta[i] = de;
    mx = fmax(mx, fmax(fmax(-1.01 * ds / xsi[i], -1.01 * de / eta[i]), fmax(-1.01 * d / xa, 1.01 * d / bx)));
  }
  mx = blk_max(mx);
  if (threadIdx.x == 0) partial[blockIdx.x] = mx;
}

// trial point at step t and its residual pieces: per constraint sum gvec_j (rows 0..m-1), then the sum of
// squares and the max |.| of the n-sized residuals (rows m, m + 1)
__global__ void k_mma_trial(long long n, int m, MmaVecs v, const double* x, const double* xsi, const double* eta,
                            const double* dx, const double* dxsi, const double* deta, double t, MmaLam lam,
                            double epsi, double* xt, double* xsit, double* etat, double* partial, int nb) {
  double sg[MMA_MAX_M];
  for (int j = 0; j < m; ++j) sg[j] = 0.0;
  double ss = 0, mx = 0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const double xv = x[i] + t * (dx ? dx[i] : 0.0);
    const double sv = xsi[i] + t * (dxsi ? dxsi[i] : 0.0);
    const double ev = eta[i] + t * (deta ? deta[i] : 0.0);
    if (xt) {
      xt[i] = xv;
      xsit[i] = sv;
      etat[i] = ev;
    }
    const double ux = v.upp[i] - xv, xl = xv - v.low[i];
    double plam = v.p0[i], qlam = v.q0[i];
    for (int j = 0; j < m; ++j) {
      const double pj = v.pp[(long long)j * n + i], qj = v.qq[(long long)j * n + i];
      plam += lam.v[j] * pj;
      qlam += lam.v[j] * qj;
      sg[j] += pj / ux + qj / xl;
    }
    const double rex = plam / (ux * ux) - qlam / (xl * xl) - sv + ev;
    const double rexsi = sv * (xv - v.alfa[i]) - epsi;
    const double reeta = ev * (v.beta[i] - xv) - epsi;
    ss += rex * rex + rexsi * rexsi + reeta * reeta;
    mx = fmax(mx, fmax(fabs(rex), fmax(fabs(rexsi), fabs(reeta))));
  }
  for (int j = 0; j < m; ++j) {
    const double a = blk_sum(sg[j]);
    if (threadIdx.x == 0) partial[(long long)j * nb + blockIdx.x] = a;
  }
  const double b = blk_sum(ss), c = blk_max(mx);
  if (threadIdx.x == 0) {
    partial[(long long)m * nb + blockIdx.x] = b;
    partial[(long long)(m + 1) * nb + blockIdx.x] = c;
  }
}

__global__ void k_copy(const double* a, double* b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_clip01(double* x, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = fmin(fmax(x[i], 0.0), 1.0);
}

// dual variables of the subproblem (y, lam, mu, s are m-vectors; z, zet scalars)
struct MmaDual {
  std::vector<double> y, lam, mu, s;
  double z = 1.0, zet = 1.0;
};

// x = a^-1 b for a small symmetric positive definite m x m matrix (Cholesky, FP64, host)
static std::vector<double> spd_solve(std::vector<double> a, std::vector<double> b, int m) {
  for (int j = 0; j < m; ++j) {
    double d = a[j * m + j];
    for (int k = 0; k < j; ++k) d -= a[j * m + k] * a[j * m + k];
    if (!(d > 0)) throw std::runtime_error("MMA Newton system is not positive definite");
    d = std::sqrt(d);
    a[j * m + j] = d;
    for (int i = j + 1; i < m; ++i) {
      double v = a[i * m + j];
      for (int k = 0; k < j; ++k) v -= a[i * m + k] * a[j * m + k];
      a[i * m + j] = v / d;
    }
  }
  for (int i = 0; i < m; ++i) {
    for (int k = 0; k < i; ++k) b[i] -= a[i * m + k] * b[k];
    b[i] /= a[i * m + i];
  }
  for (int i = m - 1; i >= 0; --i) {
    for (int k = i + 1; k < m; ++k) b[i] -= a[k * m + i] * b[k];
    b[i] /= a[i * m + i];
  }
  return b;
}

static void mma_step(long long n, int m, int it, double* x, double* xold1, double* xold2, double* low, double* upp,
                     const double* df0, const double* g, const double* dg, double move, double scale, hipStream_t s) {
  const int nb = std::max(1, std::min(256, (int)((n + 1023) / 1024)));
  const unsigned ge = (unsigned)((n + 255) / 256);
  const int rows = 2 * m + m * (m + 1) / 2;
  std::vector<double*> buf(15, nullptr);
  for (auto& p : buf) DFMI_HIP(hipMallocAsync((void**)&p, sizeof(double) * n, s));
  double *pp = nullptr, *qq = nullptr, *partial = nullptr;
  DFMI_HIP(hipMallocAsync((void**)&pp, sizeof(double) * m * n, s));
  DFMI_HIP(hipMallocAsync((void**)&qq, sizeof(double) * m * n, s));
  DFMI_HIP(hipMallocAsync((void**)&partial, sizeof(double) * rows * nb, s));
  double *alfa = buf[0], *beta = buf[1], *p0 = buf[2], *q0 = buf[3], *xs = buf[4], *xsi = buf[5], *eta = buf[6],
         *dx = buf[7], *dxsi = buf[8], *deta = buf[9], *tx = buf[10], *tsi = buf[11], *te = buf[12];
  std::vector<double> h((size_t)rows * nb);
  auto fetch = [&](int parts) {
    DFMI_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * parts * nb, hipMemcpyDeviceToHost, s));
    DFMI_HIP(hipStreamSynchronize(s));
  };
  auto psum = [&](int row) {
    double a = 0;
    for (int i = 0; i < nb; ++i) a += h[(size_t)row * nb + i];
    return a;
  };
  auto pmax = [&](int row) {
    double a = -1e300;
    for (int i = 0; i < nb; ++i) a = std::max(a, h[(size_t)row * nb + i]);
    return a;
  };
  auto pack = [&](const std::vector<double>& v) {
    MmaLam o{};
    for (int j = 0; j < m; ++j) o.v[j] = v[j];
    return o;
  };
  hipLaunchKernelGGL(k_mma_prepare, dim3(nb), dim3(256), 0, s, n, m, it, x, xold1, xold2, low, upp, df0, dg, move,
                     scale, alfa, beta, p0, q0, pp, qq, xs, xsi, eta, partial, nb);
  fetch(m);
  std::vector<double> b(m);
  for (int j = 0; j < m; ++j) b[j] = psum(j) - scale * g[j];
  const double a0 = 1.0, c = 1e4, d = 1.0;
  MmaVecs v{low, upp, alfa, beta, p0, q0, pp, qq};
  MmaDual sc;
  sc.y.assign(m, 1.0);
  sc.lam.assign(m, 1.0);
  sc.mu.assign(m, std::max(1.0, 0.5 * c));
  sc.s.assign(m, 1.0);
  double epsi = 1.0;
  // norm and max |.| of the full KKT residual from the n-sized pieces and the dual variables
  auto residual = [&](const MmaDual& q, const std::vector<double>& gvec, double ss_n, double mx_n, double& norm,
                      double& mx) {
    double ss = ss_n;
    mx = mx_n;
    auto add = [&](double r) {
      ss += r * r;
      mx = std::max(mx, std::fabs(r));
    };
    for (int j = 0; j < m; ++j) add(c + d * q.y[j] - q.mu[j] - q.lam[j]);
    add(a0 - q.zet);
    for (int j = 0; j < m; ++j) add(gvec[j] - q.y[j] + q.s[j] - b[j]);
    for (int j = 0; j < m; ++j) add(q.mu[j] * q.y[j] - epsi);
    add(q.zet * q.z - epsi);
    for (int j = 0; j < m; ++j) add(q.lam[j] * q.s[j] - epsi);
    norm = std::sqrt(ss);
  };
  std::vector<double> gvec(m);
  while (epsi > 1e-7) {
    hipLaunchKernelGGL(k_mma_trial, dim3(nb), dim3(256), 0, s, n, m, v, xs, xsi, eta, (const double*)nullptr,
                       (const double*)nullptr, (const double*)nullptr, 0.0, pack(sc.lam), epsi, (double*)nullptr,
                       (double*)nullptr, (double*)nullptr, partial, nb);
    fetch(m + 2);
    for (int j = 0; j < m; ++j) gvec[j] = psum(j);
    double norm, rmax;
    residual(sc, gvec, psum(m), pmax(m + 1), norm, rmax);
    int inner = 0;
    while (rmax > 0.9 * epsi && inner < 200) {
      ++inner;
      hipLaunchKernelGGL(k_mma_newton, dim3(nb), dim3(256), 0, s, n, m, v, xs, xsi, eta, pack(sc.lam), epsi, partial,
                         nb);
      fetch(rows);
      std::vector<double> dely(m), diagy(m), blam(m), alam((size_t)m * m);
      for (int j = 0; j < m; ++j) {
        gvec[j] = psum(j);
        dely[j] = c + d * sc.y[j] - sc.lam[j] - epsi / sc.y[j];
        const double dellam = gvec[j] - sc.y[j] - b[j] + epsi / sc.lam[j];
        diagy[j] = d + sc.mu[j] / sc.y[j];
        blam[j] = dellam + dely[j] / diagy[j] - psum(m + j);
      }
      for (int j = 0, q = 0; j < m; ++j)
        for (int l = j; l < m; ++l, ++q) alam[(size_t)j * m + l] = alam[(size_t)l * m + j] = psum(2 * m + q);
      for (int j = 0; j < m; ++j) alam[(size_t)j * m + j] += sc.s[j] / sc.lam[j] + 1.0 / diagy[j];
      const std::vector<double> dlam = spd_solve(alam, blam, m);
      const double delz = a0 - epsi / sc.z;
      const double dz = -delz * sc.z / sc.zet;  // second block row with a = 0
      hipLaunchKernelGGL(k_mma_direction, dim3(nb), dim3(256), 0, s, n, m, v, xs, xsi, eta, pack(sc.lam), pack(dlam),
                         epsi, dx, dxsi, deta, partial);
      fetch(1);
      std::vector<double> dy(m), dmu(m), ds(m);
      double stm = pmax(0);
      for (int j = 0; j < m; ++j) {
        dy[j] = -dely[j] / diagy[j] + dlam[j] / diagy[j];
        dmu[j] = -sc.mu[j] + epsi / sc.y[j] - sc.mu[j] * dy[j] / sc.y[j];
        ds[j] = -sc.s[j] + epsi / sc.lam[j] - sc.s[j] * dlam[j] / sc.lam[j];
        stm = std::max({stm, -1.01 * dy[j] / sc.y[j], -1.01 * dlam[j] / sc.lam[j], -1.01 * dmu[j] / sc.mu[j],
                        -1.01 * ds[j] / sc.s[j]});
      }
      const double dzet = -sc.zet + epsi / sc.z - sc.zet * dz / sc.z;
      stm = std::max({stm, -1.01 * dz / sc.z, -1.01 * dzet / sc.zet});
      double step = 1.0 / std::max(stm, 1.0);
      const MmaDual old = sc;
      double newnorm = 2.0 * norm;
      int tries = 0;
      while (newnorm > norm && tries < 50) {
        ++tries;
        MmaDual t = old;
        for (int j = 0; j < m; ++j) {
          t.y[j] += step * dy[j];
          t.lam[j] += step * dlam[j];
          t.mu[j] += step * dmu[j];
          t.s[j] += step * ds[j];
        }
        t.z += step * dz;
        t.zet += step * dzet;
        hipLaunchKernelGGL(k_mma_trial, dim3(nb), dim3(256), 0, s, n, m, v, xs, xsi, eta, dx, dxsi, deta, step,
                           pack(t.lam), epsi, tx, tsi, te, partial, nb);
        fetch(m + 2);
        for (int j = 0; j < m; ++j) gvec[j] = psum(j);
        residual(t, gvec, psum(m), pmax(m + 1), newnorm, rmax);
        sc = t;
        step *= 0.5;
      }
      std::swap(xs, tx);
      std::swap(xsi, tsi);
      std::swap(eta, te);
      norm = newnorm;
    }
    epsi *= 0.1;
  }
  // xold2 <- xold1, xold1 <- x, x <- clip(subproblem solution)
  hipLaunchKernelGGL(k_copy, dim3(ge), dim3(256), 0, s, xold1, xold2, n);
  hipLaunchKernelGGL(k_copy, dim3(ge), dim3(256), 0, s, x, xold1, n);
  hipLaunchKernelGGL(k_copy, dim3(ge), dim3(256), 0, s, xs, x, n);
  hipLaunchKernelGGL(k_clip01, dim3(ge), dim3(256), 0, s, x, n);
  DFMI_HIP(hipGetLastError());
  for (auto p : buf) DFMI_HIP(hipFreeAsync(p, s));
  for (auto p : {pp, qq, partial}) DFMI_HIP(hipFreeAsync(p, s));
  DFMI_HIP(hipStreamSynchronize(s));
}

}  // namespace dfmi

using namespace dfmi;

struct dfmi_filter {
  Filter f;
};

#define DFMI_GUARD2(...)                 \
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

int dfmi_simp(const double* rho, const uint8_t* active, int64_t n, double young, double penal, double eps,
              double* moduli, double* dmoduli, void* stream) {
  DFMI_GUARD2({
    if (penal < 1 || !(eps > 0 && eps < 1)) throw std::invalid_argument("Invalid SIMP interpolation settings");
    hipLaunchKernelGGL(k_simp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, rho, active,
                       (long long)n, young, penal, eps, moduli, dmoduli);
    DFMI_HIP(hipGetLastError());
  })
}

int dfmi_project(const double* filtered, const uint8_t* flags, int64_t n, double beta, const double* eta,
                 int32_t count, int32_t robust, double* physical, double* derivative, void* stream) {
  DFMI_GUARD2({
    if (n < 1 || count < 1 || count > 3 || !eta || !filtered || !flags || !physical || !derivative)
      throw std::invalid_argument("project needs n >= 1, 1 to 3 thresholds and buffers");
    if (!(beta >= 0) || !std::isfinite(beta)) throw std::invalid_argument("beta must be finite and >= 0");
    ProjectArgs pa{};
    pa.beta = beta;
    pa.count = count;
    pa.robust = robust ? 1 : 0;
    for (int q = 0; q < count; ++q) {
      if (!(eta[q] > 0 && eta[q] < 1)) throw std::invalid_argument("projection thresholds must lie within (0, 1)");
      pa.eta[q] = eta[q];
    }
    hipLaunchKernelGGL(k_project, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, (long long)n,
                       filtered, flags, pa, physical, derivative);
    DFMI_HIP(hipGetLastError());
  })
}

int dfmi_filter_create(const int32_t shape[3], const double spacing[3], double radius, const uint8_t* allowed,
                       const uint8_t* preserve, dfmi_filter** out) {
  DFMI_GUARD2({
    auto fp = std::make_unique<dfmi_filter>();
    Filter& F = fp->f;
    F.g.nx = shape[0];
    F.g.ny = shape[1];
    F.g.nz = shape[2];
    F.n = (long long)shape[0] * shape[1] * shape[2];
    for (int a = 0; a < 3; ++a) F.spacing[a] = spacing[a];
    std::vector<int> offs;
    std::vector<double> w;
    cone_taps(spacing, radius, offs, w);
    F.g.taps = (int)w.size();
    F.nb = std::max(1, std::min(256, (int)((F.n + 1023) / 1024)));
    DFMI_HIP(hipMalloc(&F.offs, sizeof(int) * offs.size() + 4));
    DFMI_HIP(hipMalloc(&F.w, sizeof(double) * w.size() + 8));
    DFMI_HIP(hipMemcpy(F.offs, offs.data(), sizeof(int) * offs.size(), hipMemcpyHostToDevice));
    DFMI_HIP(hipMemcpy(F.w, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    std::vector<uint8_t> fr(F.n);
    for (long long i = 0; i < F.n; ++i) fr[i] = allowed[i] && !(preserve && preserve[i]);
    for (void** p : {&F.allowed, &F.free_}) DFMI_HIP(hipMalloc(p, F.n));
    DFMI_HIP(hipMemcpy(F.allowed, allowed, F.n, hipMemcpyHostToDevice));
    DFMI_HIP(hipMemcpy(F.free_, fr.data(), F.n, hipMemcpyHostToDevice));
    for (void** p : {&F.sums, &F.inv, &F.tmp, &F.tmp2, &F.lo, &F.hi, &F.ratio, &F.prop})
      DFMI_HIP(hipMalloc(p, sizeof(double) * F.n));
    DFMI_HIP(hipMalloc(&F.partial, sizeof(double) * 3 * F.nb));
    // sums = (W a) a, 1 on forbidden cells
    std::vector<double> ones(F.n, 1.0);
    DFMI_HIP(hipMemcpy(F.tmp, ones.data(), sizeof(double) * F.n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_cone, dim3((unsigned)((F.n + 255) / 256)), dim3(256), 0, nullptr, F.g, (const int*)F.offs,
                       (const double*)F.w, (const uint8_t*)F.allowed, (const double*)F.tmp, nullptr, nullptr,
                       nullptr, (double*)F.sums);
    int* bad = nullptr;
    DFMI_HIP(hipMalloc((void**)&bad, sizeof(int)));
    DFMI_HIP(hipMemset(bad, 0, sizeof(int)));
    hipLaunchKernelGGL(k_sums_fix, dim3((unsigned)((F.n + 255) / 256)), dim3(256), 0, nullptr,
                       (const uint8_t*)F.allowed, (double*)F.sums, F.n, bad);
    hipLaunchKernelGGL(k_inverse, dim3((unsigned)((F.n + 255) / 256)), dim3(256), 0, nullptr, (const double*)F.sums,
                       (double*)F.inv, F.n);
    int hb = 0;
    DFMI_HIP(hipMemcpy(&hb, bad, sizeof(int), hipMemcpyDeviceToHost));
    (void)hipFree(bad);
    if (hb) throw std::invalid_argument("Density filter contains an empty allowed-cell neighborhood");
    *out = fp.release();
  })
}

int dfmi_filter_destroy(dfmi_filter* f) { DFMI_GUARD2({ delete f; }) }

int dfmi_filter_set_stream(dfmi_filter* f, void* stream) { DFMI_GUARD2({ f->f.stream = (hipStream_t)stream; }) }

int dfmi_filter_forward(dfmi_filter* f, const double* x, double* out) {
  DFMI_GUARD2({
    f->f.forward(x, out);
    DFMI_HIP(hipStreamSynchronize(f->f.stream));
  })
}

int dfmi_filter_pullback(dfmi_filter* f, const double* sens, const double* dproj, double* out) {
  DFMI_GUARD2({
    Filter& F = f->f;
    F.pullback(sens, dproj, out);
    DFMI_HIP(hipStreamSynchronize(F.stream));
  })
}

int dfmi_filter_set_corridor(dfmi_filter* f, double radius, const double* weights) {
  DFMI_GUARD2({
    Filter& F = f->f;
    if (!weights) throw std::invalid_argument("corridor weights missing");
    if (!(radius > 0)) throw std::invalid_argument("corridor radius must be positive");
    std::vector<uint8_t> allowed(F.n);
    DFMI_HIP(hipMemcpy(allowed.data(), F.allowed, F.n, hipMemcpyDeviceToHost));
    std::vector<double> blend(F.n);
    std::vector<uint8_t> support(F.n);
    bool any = false;
    for (long long i = 0; i < F.n; ++i) {
      blend[i] = allowed[i] ? std::min(std::max(weights[i], 0.0), 1.0) : 0.0;
      support[i] = blend[i] > 0;
      any = any || support[i];
    }
    F.corridor = false;
    if (!any) return 0;  // no allowed cell carries a corridor weight: the plain filter stays
    std::vector<int> offs;
    std::vector<double> w;
    cone_taps(F.spacing, radius, offs, w);
    F.g2 = F.g;
    F.g2.taps = (int)w.size();
    for (void** p : {&F.offs2, &F.w2, &F.blend, &F.support, &F.inv2, &F.tmp3})
      if (*p) (void)hipFree(*p), *p = nullptr;
    DFMI_HIP(hipMalloc(&F.offs2, sizeof(int) * offs.size() + 4));
    DFMI_HIP(hipMalloc(&F.w2, sizeof(double) * w.size() + 8));
    DFMI_HIP(hipMalloc(&F.support, F.n));
    for (void** p : {&F.blend, &F.inv2, &F.tmp3}) DFMI_HIP(hipMalloc(p, sizeof(double) * F.n));
    DFMI_HIP(hipMemcpy(F.offs2, offs.data(), sizeof(int) * offs.size(), hipMemcpyHostToDevice));
    DFMI_HIP(hipMemcpy(F.w2, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    DFMI_HIP(hipMemcpy(F.blend, blend.data(), sizeof(double) * F.n, hipMemcpyHostToDevice));
    DFMI_HIP(hipMemcpy(F.support, support.data(), F.n, hipMemcpyHostToDevice));
    // wide sums (W_wide a) a on the corridor support, 1 elsewhere
    std::vector<double> ones(F.n, 1.0), sums(F.n);
    DFMI_HIP(hipMemcpy(F.tmp3, ones.data(), sizeof(double) * F.n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_cone, dim3((unsigned)((F.n + 255) / 256)), dim3(256), 0, nullptr, F.g2, (const int*)F.offs2,
                       (const double*)F.w2, (const uint8_t*)F.allowed, (const double*)F.tmp3, nullptr, nullptr,
                       (const uint8_t*)F.support, (double*)F.inv2);
    DFMI_HIP(hipGetLastError());
    DFMI_HIP(hipMemcpy(sums.data(), F.inv2, sizeof(double) * F.n, hipMemcpyDeviceToHost));
    for (long long i = 0; i < F.n; ++i) {
      if (!support[i]) sums[i] = 1.0;
      if (!(sums[i] > 0)) throw std::runtime_error("corridor filter has an empty allowed-cell neighbourhood");
      sums[i] = 1.0 / sums[i];
    }
    DFMI_HIP(hipMemcpy(F.inv2, sums.data(), sizeof(double) * F.n, hipMemcpyHostToDevice));
    F.corridor = true;
  })
}

int dfmi_filter_volume(dfmi_filter* f, const double* x, double* volume) {
  DFMI_GUARD2({
    Filter& F = f->f;
    F.forward(x, (double*)F.tmp2);
    *volume = F.sum((const double*)F.tmp2);
  })
}

int dfmi_oc_update(dfmi_filter* f, const double* x, const double* dc, const double* dv, double target, double xmin,
                   double move, double* out) {
  DFMI_GUARD2({
    Filter& F = f->f;
    hipStream_t s = F.stream;
    const unsigned ge = (unsigned)((F.n + 255) / 256);
    const uint8_t* fr = (const uint8_t*)F.free_;
    hipLaunchKernelGGL(k_oc_check, dim3(F.nb), dim3(256), 0, s, dc, dv, fr, F.n, (double*)F.partial, F.nb);
    std::vector<double> h(3 * F.nb);
    DFMI_HIP(hipMemcpyAsync(h.data(), F.partial, sizeof(double) * 3 * F.nb, hipMemcpyDeviceToHost, s));
    DFMI_HIP(hipStreamSynchronize(s));
    double mn = 1e300, mx = -1e300, bad = 0;
    for (int i = 0; i < F.nb; ++i) {
      mn = std::min(mn, h[i]);
      mx = std::max(mx, h[F.nb + i]);
      bad += h[2 * F.nb + i];
    }
    if (mn < 0 || !(mx > 0) || bad > 0) throw std::runtime_error("Invalid OC sensitivities");
    double *ratio = (double*)F.ratio, *lo = (double*)F.lo, *hi = (double*)F.hi, *prop = (double*)F.prop;
    hipLaunchKernelGGL(k_oc_prep, dim3(ge), dim3(256), 0, s, x, dc, dv, fr, mx, xmin, move, ratio, lo, hi, F.n);
    auto volume_of = [&](double lmb, int lowest, double* dst) {
      hipLaunchKernelGGL(k_oc_prop, dim3(ge), dim3(256), 0, s, x, fr, ratio, lo, hi, lmb, lowest, dst, F.n);
      F.forward(dst, (double*)F.tmp2);
      return F.sum((const double*)F.tmp2);
    };
    if (volume_of(0.0, 1, out) > target) return 0;
    hipLaunchKernelGGL(k_ratio_max, dim3(F.nb), dim3(256), 0, s, ratio, fr, F.n, (double*)F.partial);
    DFMI_HIP(hipMemcpyAsync(h.data(), F.partial, sizeof(double) * F.nb, hipMemcpyDeviceToHost, s));
    DFMI_HIP(hipStreamSynchronize(s));
    double rmax = 0;
    for (int i = 0; i < F.nb; ++i) rmax = std::max(rmax, h[i]);
    double upper = std::max(rmax, 1e-12);
    bool ok = false;
    for (int t = 0; t < 100; ++t) {
      if (volume_of(upper, 0, prop) <= target + 1e-9) {
        ok = true;
        break;
      }
      upper *= 2;
    }
    if (!ok) throw std::runtime_error("OC could not bracket the volume multiplier");
    double lower = 0.0;
    volume_of(upper, 0, out);
    for (int t = 0; t < 70; ++t) {
      const double mid = 0.5 * (lower + upper);
      if (volume_of(mid, 0, prop) > target) {
        lower = mid;
      } else {
        upper = mid;
        hipLaunchKernelGGL(k_copy, dim3(ge), dim3(256), 0, s, (const double*)prop, out, F.n);
      }
      if ((upper - lower) / std::max(upper + lower, 1e-100) < 1e-8) break;
    }
    DFMI_HIP(hipStreamSynchronize(s));
  })
}

int dfmi_mma_step(int64_t n, int32_t m, int32_t iteration, double* x, double* xold1, double* xold2, double* low,
                  double* upp, double f0, const double* df0, const double* g, const double* dg, double move,
                  double scale, void* stream) {
  DFMI_GUARD2({
    (void)f0;  // the subproblem does not use the objective value (mmasub only passes it through)
    if (n < 1) throw std::invalid_argument("MMA needs n >= 1");
    if (m < 1 || m > MMA_MAX_M) throw std::invalid_argument("MMA takes 1 to 16 constraints");
    if (!x || !xold1 || !xold2 || !low || !upp || !df0 || !g || !dg) throw std::invalid_argument("null MMA argument");
    mma_step((long long)n, m, iteration, x, xold1, xold2, low, upp, df0, g, dg, move, scale, (hipStream_t)stream);
  })
}

}  // extern "C"
